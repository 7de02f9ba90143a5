/*
 * bmg.h -- C ABI of the MI355X geometric-multigrid Poisson solver.
 *
 * Drop-in for Briscola's MG linear solver:
 *   template<class SType,class Type,class MeshType> class MG
 *       : public solver<SType,Type,MeshType>
 *   virtual void solve(linearSystem<SType,Type,MeshType>& sys,
 *                      const bool constMatrix = false);
 * (src/briscolaFiniteVolume/solvers/MG/MG.H:135-139, MG.C:402-483; the
 * abstract interface is solvers/solver/solver.H:280-284).  The hot
 * instantiation is <stencil, scalar, colocated>.
 *
 * Layout at the boundary is the reference's own (no conversion needed by the
 * caller):
 *   - a level field is a ghost-padded block of (Nx+2)(Ny+2)(Nz+2) doubles,
 *     lin(i,j,k) = (i*(Ny+2)+j)*(Nz+2)+k, k (z) contiguous
 *     (src/briscolaCore/include/utils.H:559-570, block.H:37-393);
 *   - a stencil field stores 7 doubles [C,L,R,B,T,A,F] per padded cell
 *     (src/briscolaCore/primitives/StencilSpace/stencil/stencil.H:21).
 *
 * Conventions: every entry point returns BMG_OK (0) or a negative status;
 * bmg_last_error() gives the message of the last failure on the calling
 * thread.  No exception crosses the ABI.  A context owns its device memory
 * and one HIP stream; distinct contexts may be used from different threads,
 * one context may not.
 */
#ifndef BMG_H
#define BMG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bmg_ctx bmg_ctx;

enum bmg_status {
    BMG_OK = 0,
    BMG_ERR_ARG = -1,        /* invalid argument / shape */
    BMG_ERR_HIP = -2,        /* HIP runtime error */
    BMG_ERR_STATE = -3,      /* call out of order (e.g. no operator) */
    BMG_ERR_SINGULAR = -4,   /* coarse matrix could not be factorised */
    BMG_ERR_NUMERIC = -5     /* NaN/Inf residual norm */
};

/* Boundary kinds per face, order -x,+x,-y,+y,-z,+z (faceOffsets,
 * src/briscolaCore/include/utils.H:191-241). */
enum bmg_bc { BMG_NEUMANN = 0, BMG_DIRICHLET = 1, BMG_PERIODIC = 2 };
/* MGCycleType, MG.H:18-23 */
enum bmg_cycle { BMG_VCYCLE = 0, BMG_WCYCLE = 1, BMG_FCYCLE = 2 };
/* MGCoarseMode, MG.H:26-30 */
enum bmg_coarse_mode { BMG_COARSE_SMOOTH = 0, BMG_COARSE_DIRECT = 1 };
/* smoother TypeNames rbgs / jac / lexgs (rbgsSmoother.H:56, jacSmoother.H,
 * lexgsSmoother.H).  BMG_LEXGS keeps the sequential lexicographic order
 * exactly (one launch per hyperplane i+j+k = const of the swept region), so
 * it is far slower than BMG_RBGS; it is there for parity with setups that
 * name it. */
enum bmg_smoother { BMG_RBGS = 0, BMG_JAC = 1, BMG_LEXGS = 2 };
/* BMG_FP64: every pass in FP64 (the reference's arithmetic).
 * BMG_MIXED: FP32 cycles on the defect equation inside FP64 defect
 * correction; FP64 residual, update and coarsest direct solve. */
enum bmg_precision { BMG_FP64 = 0, BMG_MIXED = 1 };
/* Outer iteration: BMG_KRYLOV_NONE repeats MG cycles as MG::Solve does
 * (MG.C:191-242); BMG_KRYLOV_FCG runs flexible conjugate gradients (Notay,
 * FCG with one stored direction) preconditioned by one MG cycle of the
 * chosen precision, with the same stopping test; BMG_KRYLOV_FCG_LEARNED
 * starts FCG from the guess of the attached learned network and
 * preconditions with it until an iteration reduces the residual by less
 * than a factor 0.3, then with MG cycles (bmg_unet.h). */
enum bmg_krylov { BMG_KRYLOV_NONE = 0, BMG_KRYLOV_FCG = 1,
                  BMG_KRYLOV_FCG_LEARNED = 2 };

/* Mesh + boundary conditions: what MG reads from fvMesh and the BCs of
 * sys.x() (mesh.C:19-53 hierarchy, fvMeshMetrics.C metrics,
 * Neumann/Dirichlet/periodicBoundaryCondition.C). */
typedef struct bmg_grid {
    int n[3];                   /* level-0 cells per axis (i,j,k) = (x,y,z) */
    const double* vertices[3];  /* n[d]+1 vertex coordinates per axis       */
    int bc[6];                  /* enum bmg_bc per face                     */
    double bc_value[6];         /* Neumann gradient / Dirichlet value,
                                   uniform per face, level 0 only (per-cell
                                   data: bmg_set_boundary_values)           */
} bmg_grid;

/* The MG solver dictionary (MG.C:257-304, solver.C:114-123). */
typedef struct bmg_opts {
    double rel_tol;         /* relTol (required)            */
    double tolerance;       /* tolerance (required)         */
    int min_iter;           /* minIter, default 0           */
    int max_iter;           /* maxIter, default 999         */
    int n_sweeps_pre;       /* nSweepsPre, default 0        */
    int n_sweeps_post;      /* nSweepsPost, default 2       */
    int cycle;              /* cycleType, default F         */
    int coarse_mode;        /* coarseMode, default direct   */
    int coarse_level;       /* coarseLevel, default 0       */
    int smoother;           /* smoother, default rbgs       */
    int precision;          /* BMG_FP64 (default) / BMG_MIXED */
    int krylov;             /* BMG_KRYLOV_NONE (default) / BMG_KRYLOV_FCG */
    int deflation;          /* 0 (default) / 1: flexible CG deflated by the
                               connected sets of high-coefficient cells
                               (see below)                          */
    int reserved[3];
} bmg_opts;
/* Deflation (opts.deflation = 1, BMG_KRYLOV_FCG only; no reference
 * counterpart): when bmg_set_operator_laplacian assembles an operator whose
 * cell coefficients (the mean of a cell's six face lambdas) span a ratio of
 * at least 100, the cells above the geometric mean of the extremes are
 * grouped into face-connected sets (across periodic faces too), and the
 * largest sets of at least 8 cells, at most 64, become indicator vectors
 * Z.  Flexible CG then starts from x0 + Z E^-1 Z^T r0 (E = Z^T A Z) and
 * projects every new search direction p to p - Z E^-1 (A Z)^T p (in FP64,
 * the same as projecting the preconditioned residual): the
 * near-constant modes of high-coefficient inclusions (bubbles, droplets),
 * which geometric coarsening cannot represent, are removed exactly.  The
 * stopping test is unchanged.  Other operators or options solve as without
 * deflation; bmg_deflation_vectors reports the number of vectors. */

typedef struct bmg_stats {
    int iters;              /* nIter of printSolverStats (solver.C:97-112) */
    int status;             /* BMG_OK or the failing status               */
    double initial_residual;
    double final_residual;
    double solve_ms;        /* device time of the whole solve             */
    int learned_iters;      /* FCG iterations preconditioned by the network */
    int learned_fallback;   /* iteration after which MG cycles took over
                               (-1: never)                                 */
    double learned_guess_residual;  /* residual after the network's initial
                                       guess (normalised; 0: no guess)     */
    double reserved[2];
} bmg_stats;

/* Number of deflation vectors of the current operator (0: none). */
int bmg_deflation_vectors(const bmg_ctx* ctx, int* count);

/* Fill opts with the reference defaults (rel_tol/tolerance set to 0). */
void bmg_default_opts(bmg_opts* opts);

const char* bmg_last_error(void);
int bmg_device_count(int* count);

/* Builds the level hierarchy (mesh::generateLevels, mesh.C:19-53; coarsen,
 * utils.H:1154-1171), level metrics and transfer weights
 * (linearRestrictionScheme.C:15-80, linearProlongationScheme.C:12-110) and
 * allocates device storage on the current HIP device. */
int bmg_create(bmg_ctx** ctx, const bmg_grid* grid, const bmg_opts* opts);
int bmg_destroy(bmg_ctx* ctx);
int bmg_set_opts(bmg_ctx* ctx, const bmg_opts* opts);

/* Per-cell boundary data of one Neumann (gradient) or Dirichlet (value)
 * face on level 0, replacing the face's uniform bc_value: the patch block
 * of the face's boundary condition, boundaryValues()[0] of a Dirichlet
 * condition (DirichletBoundaryConditionBase.H:30,82) or
 * boundaryGradients()[0] of a Neumann one (NeumannBoundaryConditionBase.H:
 * 30,82), read cell by cell as boundaryValues_[d](ijk - S)
 * (DirichletBoundaryCondition.C:45,68, NeumannBoundaryCondition.C:51,83).
 * Layout: the face's boundary layer of level-0 interior cells with the
 * normal axis of size 1, (i,j,k) order with the last axis fastest -- for
 * face -x or +x, Ny*Nz values, value (j,k) at j*Nz + k; for -y/+y, Nx*Nz
 * values at i*Nz + k; for -z/+z, Nx*Ny values at i*Ny + j.  The data is
 * copied; it enters b through the ghost elimination of the next solve after
 * the operator is set (as eliminateGhosts folds it, MG.C:445) and the ghost
 * values written at the end of every solve (MG.C:475).  values NULL: back
 * to the uniform value.  BMG_ERR_ARG for a periodic face. */
int bmg_set_boundary_values(bmg_ctx* ctx, int face, const double* values);
/* Same, device memory. */
int bmg_set_boundary_values_device(bmg_ctx* ctx, int face,
                                   const double* d_values);

int bmg_num_levels(const bmg_ctx* ctx);
int bmg_level_shape(const bmg_ctx* ctx, int level, int n[3]);

/* Operator input, drop-in form: sys.A()[level] as assembled (before ghost
 * elimination), reference stencil layout, host memory.  Every level must be
 * set before a solve.  symmetric: linearSystem::symmetric_ (decides
 * singular(), linearSystem.H:227-230). */
int bmg_set_operator_stencil(bmg_ctx* ctx, int level, const double* A7,
                             int symmetric);
/* Same, device memory. */
int bmg_set_operator_stencil_device(bmg_ctx* ctx, int level,
                                    const double* d_A7, int symmetric);

/* Operator input, native form: the solver assembles
 * laplacian(lambda, x) = div(lambda grad x) itself on every level
 * (linearGaussLaplacianScheme.C:22-102).  lambda: NULL for the lambda-free
 * operator, else three host arrays of level-0 lower-face coefficients in
 * padded cell layout, restricted to coarse levels area-weighted
 * (areaWeightedFaceRestrictionScheme.C:19-56). */
int bmg_set_operator_laplacian(bmg_ctx* ctx, const double* lambda_x,
                               const double* lambda_y,
                               const double* lambda_z);
/* The same minus the implicit Euler time term, the system defaultPoissonSolver
 * builds with ddt = true (defaultPoissonSolver.C:86-94): sys -= im::ddt(x) /
 * dt_frac, whose matrix is the diagonal V/delta_t of every level
 * (EulerDdtScheme.C:50-54; V = (dx*dy)*dz of the level's cell), so every
 * centre coefficient is lowered by (V/delta_t)/dt_frac before ghost
 * elimination.  The matching source term, b_sys -= ((V*x_old)/delta_t)/
 * dt_frac on level 0, is the caller's part of b.  delta_t > 0 and a finite
 * nonzero dt_frac; BMG_ERR_ARG otherwise.  Replaces
 * defaultPoissonSolver.C:91-92 with Euler as the ddt scheme. */
int bmg_set_operator_laplacian_ddt(bmg_ctx* ctx, const double* lambda_x,
                                   const double* lambda_y,
                                   const double* lambda_z, double delta_t,
                                   double dt_frac);

/* The operator as last set stands for a newly assembled linearSystem
 * again: the next solve folds the level-0 boundary terms of the Dirichlet
 * values and Neumann gradients into its b, as MG::solve's eliminateGhosts
 * does for a system whose ghost coefficients are still in place
 * (MG.C:445, DirichletBoundaryCondition.C:36-46,
 * NeumannBoundaryCondition.C:40-54), from the coefficients before
 * elimination, without assembling or eliminating the operator again (the
 * coarse factorisation stays valid for const_matrix).  Setting an operator
 * implies it.  For callers that build a new system with the same matrix
 * for every solve (defaultPoissonSolver.C:39-124 with constMatrix); a
 * solve without it after a previous one treats b as the caller's
 * already-eliminated system, as the reference does when the same mutated
 * linearSystem is solved again. */
int bmg_renew_source_terms(bmg_ctx* ctx);

/* Read back sys.A()[level] after the solver's ghost elimination and
 * singularity pin (the reference mutates the caller's A, MG.C:433-447). */
int bmg_get_operator_stencil(bmg_ctx* ctx, int level, double* A7);

/* MG::solve: b (level-0 source, in/out: inhomogeneous boundary terms are
 * folded in as eliminateGhosts does) and x (initial guess in, solution out,
 * eliminated ghosts refreshed on exit, MG.C:475).  Host pointers, reference
 * layout.  const_matrix: reuse the coarse factorisation (MG.C:449-459). */
int bmg_solve(bmg_ctx* ctx, double* b, double* x, int const_matrix,
              bmg_stats* stats);
/* Same with device pointers on the context's device; runs on the context's
 * stream and returns after the solve completed. */
int bmg_solve_device(bmg_ctx* ctx, double* d_b, double* d_x,
                     int const_matrix, bmg_stats* stats);

/* ---- Batched solves of independent systems.  Each system is a context
 * created and given its operator as for bmg_solve_device; the batch runs
 * MG::solve (MG.C:402-483; rbgs smoother) of all of them together, with MG
 * iterations (krylov none) or flexible CG (BMG_KRYLOV_FCG, the same for
 * every system; the learned preconditioner has no batched form): every
 * kernel of an iteration is one persistent launch over all systems, the
 * merged iteration is replayed as a HIP graph, and converged systems drop
 * out of it.  Each system's result is bitwise the one bmg_solve_device
 * returns for it alone.  The contexts stay owned by the caller, must be on
 * one device and must not be used elsewhere while the batch solves; they
 * may differ in shape, boundary conditions, precision and cycle options.
 * No reference counterpart: the reference solves one system per call
 * (solver.H:280-284). ---- */
typedef struct bmg_batch bmg_batch;
int bmg_batch_create(bmg_batch** batch, bmg_ctx* const* ctxs, int n);
int bmg_batch_destroy(bmg_batch* batch);
/* d_b[i], d_x[i]: device pointers of system i as in bmg_solve_device;
 * stats: n entries (solve_ms = the whole batch). */
int bmg_solve_batch(bmg_batch* batch, double* const* d_b, double* const* d_x,
                    int const_matrix, bmg_stats* stats);
/* Launches and virtual workgroups of one merged iteration of the last
 * solve's outer iteration (0 before the first solve). */
int bmg_batch_info(const bmg_batch* batch, int* launches,
                   long* workgroups);
/* Autotuning of the merged iteration (default on): every launch of a
 * kernel with several tile shapes (red-black phases, residuals without a
 * norm, restriction, prolongation) first takes the fastest shape for all
 * its shares -- rows of 64, 32, 16 or 8 cells per wavefront row, or the
 * shapes recorded (the row width with the fewest idle lanes) -- then every
 * launch the fastest persistent grid of 4, 8, 16 or 32 workgroups per CU;
 * each by timing the whole sequence.  Off: recorded shapes, 16 workgroups
 * per CU.  Rebuilds the plan.  The per-cell arithmetic, and so every
 * result, is the same under every shape and grid. */
int bmg_batch_set_autotune(bmg_batch* batch, int enable);
/* Fixes the tile shape of every launch that takes several (ws 3..6: rows
 * of 2^ws cells per wavefront row) instead of letting the autotuner pick
 * it; 0 (default): picked by timing.  Rebuilds the plan. */
int bmg_batch_set_tile_shape(bmg_batch* batch, int ws);
/* Launches of the last solve's merged iteration and how many of them run
 * another tile shape than recorded. */
int bmg_batch_tuning(const bmg_batch* batch, int* launches, int* reshaped);

/* ---- FFT Poisson solver (the reference's FFTPoissonSolver,
 * FFTPoissonSolver.C:97-465): div(grad x) = -b in finite-difference form on
 * a rectilinear grid with at least two uniform directions, by FFTs over
 * the two transformed directions and a tridiagonal solve per wave-number
 * pair along the other one (the first non-uniform direction, else x;
 * planner.C:76-117).  Transformed directions may be periodic,
 * Neumann-Neumann or Dirichlet-Dirichlet; the tridiagonal direction takes
 * any pair.  b: the source (null: zero); x: in, its ghosts give the
 * boundary terms (FFTPoissonSolver.C:123-180); out, the solution with every
 * ghost refreshed.  Reference layout.  With only Neumann and periodic faces
 * the zero mode is pinned as the reference pins it (the solution is defined
 * up to a constant).  Runs on the current device and its own stream. ---- */
typedef struct bmg_fft bmg_fft;
int bmg_fft_create(bmg_fft** fft, const bmg_grid* grid);
int bmg_fft_destroy(bmg_fft* fft);
int bmg_fft_solve(bmg_fft* fft, const double* b, double* x);
int bmg_fft_solve_device(bmg_fft* fft, const double* d_b, double* d_x);
/* The same with the implicit time term, as the reference's solve(x, b,
 * nullptr, ddt = true, dtFrac) (FFTPoissonSolver.C:116-131,
 * tridiagonalSolver.C:334-352, :430): with dT = delta_t*dt_frac the source
 * becomes -b - x_old/dT and every diagonal of the tridiagonal systems is
 * lowered by 1/dT (the pinned zero-mode row of an all Neumann/periodic
 * mesh included, as the reference does).  x_old: the old-time field,
 * reference layout.  BMG_ERR_ARG without x_old or with dT <= 0. */
int bmg_fft_solve_ddt(bmg_fft* fft, const double* b, double* x,
                      const double* x_old, double delta_t, double dt_frac);
int bmg_fft_solve_ddt_device(bmg_fft* fft, const double* d_b, double* d_x,
                             const double* d_x_old, double delta_t,
                             double dt_frac);

/* Correction source of the split Poisson solver (splitPoissonSolver.C:
 * 124-137) on level 0 of ctx's mesh: per face direction d, on the faces of
 * bmg_face_flux's layout, phi_d = (((lambda_d/lambda0) - 1) * g) * A_d with
 * the face gradient g = (x[c - e_d] - x[c]) * delta_d and face area A_d;
 * then per interior cell bHat = b/lambda0 + div(phi), the divergence
 * accumulated as divergenceScheme.C:66-95 does (lower face added, upper
 * face subtracted, directions in order) times 1/((dx*dy)*dz).  b may be
 * null (zero).  lambda0 > 0 (the reference takes the largest face
 * coefficient).  phi and bHat are written on those faces and interior
 * cells; their other entries keep what they held.  Replaces the source assembly of
 * splitPoissonSolver::solve; the caller runs the inner solve on bHat. */
int bmg_split_source(bmg_ctx* ctx, const double* x, const double* lambda_x,
                     const double* lambda_y, const double* lambda_z,
                     double lambda0, const double* b, double* bhat,
                     double* phi_x, double* phi_y, double* phi_z);
int bmg_split_source_device(bmg_ctx* ctx, const double* d_x,
                            const double* d_lambda_x, const double* d_lambda_y,
                            const double* d_lambda_z, double lambda0,
                            const double* d_b, double* d_bhat, double* d_phi_x,
                            double* d_phi_y, double* d_phi_z);

/* Flux of a level-0 field through every face, as PoissonSolver computes it
 * after a solve (defaultPoissonSolver.C:108-121): for axis d and padded cell
 * c whose lower face lies in the face range (1 <= c_d <= N_d + 1, interior
 * along the other axes), flux_d[c] = (lambda_d[c] * ((x[c - e_d] - x[c]) *
 * delta_d)) * A_d -- the face gradient along the outward normal of the lower
 * face (linearFaceGradientScheme.C:44-53) with delta_d = 1/|cc_c -
 * cc_{c-e_d}|, times the face coefficient, times the face area.  lambda
 * arrays (the padded lower-face layout of bmg_set_operator_laplacian) may
 * all be null (lambda = 1, no product).  Reference layout: bmg_face_flux
 * takes host arrays and sets every other flux entry to 0;
 * bmg_face_flux_device takes device pointers and leaves other entries as
 * they are.  No counterpart in the MG solver itself. */
int bmg_face_flux(bmg_ctx* ctx, const double* x, const double* lambda_x,
                  const double* lambda_y, const double* lambda_z,
                  double* flux_x, double* flux_y, double* flux_z);
int bmg_face_flux_device(bmg_ctx* ctx, const double* d_x,
                         const double* d_lambda_x, const double* d_lambda_y,
                         const double* d_lambda_z, double* d_flux_x,
                         double* d_flux_y, double* d_flux_z);

/* bmg_face_flux without allocation: the same host arrays, the same results,
 * bit for bit, for callers that cannot allocate device memory (code run
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  The inputs and fluxes are staged in a
 * fixed device array of BMG_NOALLOC_SCRATCH_DOUBLES doubles (32 MiB) in the
 * library's code object, one per device, held until the process ends; the
 * call makes no hipMalloc/hipFree.  Level 0 must fit: 7 x padded cells
 * doubles with lambda arrays (up to 599 186 padded cells, e.g. 82^3), 4 x
 * without; larger levels get BMG_ERR_ARG.  Every call on a device shares
 * that device's array: calls that overlap on one device (from several
 * threads or contexts) race and must be serialised by the caller. */
#define BMG_NOALLOC_SCRATCH_DOUBLES 4194304
int bmg_face_flux_noalloc(bmg_ctx* ctx, const double* x,
                          const double* lambda_x, const double* lambda_y,
                          const double* lambda_z, double* flux_x,
                          double* flux_y, double* flux_z);

/* MG::solve of a diagonal system (sys.diagonal(), MG.C:404-428): level-0
 * stencil A7 (reference layout, only the centre is used after ghost
 * elimination), x = b / A_C on every interior cell, then the level-0
 * boundary update.  b and x as in bmg_solve. */
int bmg_solve_diagonal(bmg_ctx* ctx, const double* A7, double* b, double* x);

/* ---- Level operations for kernel-level parity tests.  Fields are host
 * arrays in reference layout; 'which' is 0 = x, 1 = b, 2 = r. ---- */
int bmg_upload_field(bmg_ctx* ctx, int which, int level, const double* v);
int bmg_download_field(bmg_ctx* ctx, int which, int level, double* v);
/* MG.C:431-447 setSingularityConstraint + eliminateGhosts (+ coarse LU). */
int bmg_op_prepare(bmg_ctx* ctx);
/* rbgsSmoother::Smooth / jacSmoother::Smooth (rbgsSmoother.C:376-437). */
int bmg_op_smooth(bmg_ctx* ctx, int level, int sweeps);
/* linearSystem::residual (linearSystemResidual.C:280-332). */
int bmg_op_residual(bmg_ctx* ctx, int level, int corr_bcs);
/* b[level+1] = scaled linear restriction of r[level]
 * (linearRestrictionScheme.C:103-160). */
int bmg_op_restrict(bmg_ctx* ctx, int level);
/* x[level] += linear prolongation of x[level+1]
 * (linearProlongationScheme.C:112-147). */
int bmg_op_prolong(bmg_ctx* ctx, int level);
/* meshLevel::correctBoundaryConditions on x[level] (meshLevel.C:231-254). */
int bmg_op_correct_bcs(bmg_ctx* ctx, int level);
/* APLU::solve on the coarse level (APLU.C:91-128). */
int bmg_op_coarse_solve(bmg_ctx* ctx);
/* One MG::cycle from level 0 (MG.C:19-129). */
int bmg_op_cycle(bmg_ctx* ctx);
/* sum of r[0]^2 over interior cells. */
int bmg_op_sum_sq_residual(bmg_ctx* ctx, double* out);

/* ---- Measurement hooks (no counterpart in the reference). ----
 * bmg_time_smoother: launches the red-black phase kernel of 'level' (the
 * interior region, colours alternating) 'reps' times between two HIP events
 * on the context's stream and returns the average launch duration and the
 * algorithmic HBM bytes of one launch (cells updated x bytes per update of
 * the operator representation: b, the other colour of x, the write of x,
 * plus stored coefficients).  precision: BMG_FP64 or BMG_MIXED (FP32 arrays).
 * Needs bmg_op_prepare or a solve first; overwrites x of that level. */
int bmg_time_smoother(bmg_ctx* ctx, int level, int precision, int reps,
                      double* avg_ms, double* bytes);
/* bmg_time_op: the same for one operation of a level, launched 'reps' times
 * back to back between two HIP events on the context's stream (after one
 * untimed launch):
 *   BMG_TIME_PHASE      one red-black phase over the interior (as above);
 *   BMG_TIME_RESIDUAL   the residual of the level (FP64: with the fused norm
 *                       partials, as the iteration's level-0 residual);
 *   BMG_TIME_RESTRICT   restriction of the level's residual to level+1;
 *   BMG_TIME_PROLONG    prolongation-add of level+1 into the level (without
 *                       the coarse ghost refresh);
 *   BMG_TIME_ITERATION  one whole MG iteration (level must be 0): the
 *                       captured graph of cycle, level-0 residual and norm.
 * bytes: algorithmic HBM bytes of one launch (every array touched once:
 * phase n/2 (3w + coefficients), residual n (3w + coefficients),
 * restriction n w + n_c w, prolongation 2 n w + n_c w, w = 8 FP64 / 4 FP32,
 * coefficients 0 matrix-free, 6w face, 7w stored stencil; for an iteration
 * the sum over its launches).  Needs bmg_op_prepare or a solve first;
 * overwrites the fields of the level (an iteration: of every level). */
enum
{
    BMG_TIME_PHASE = 0,
    BMG_TIME_RESIDUAL = 1,
    BMG_TIME_RESTRICT = 2,
    BMG_TIME_PROLONG = 3,
    BMG_TIME_ITERATION = 4
};
int bmg_time_op(bmg_ctx* ctx, int op, int level, int precision, int reps,
                double* avg_ms, double* bytes);
/* Replay each MG iteration as a captured HIP graph (default on; the
 * environment variable BMG_NO_GRAPHS=1 turns it off at creation). */
int bmg_set_graphs(bmg_ctx* ctx, int enable);
/* Smoothing on levels without periodic faces can run consecutive red-black
 * phases in pairs, one launch per pair (an LDS-staged two-phase kernel that
 * reads each coefficient and right-hand-side value once per pair; the same
 * per-cell arithmetic, so the same bits as one launch per phase).
 * mode 0 (default): one launch per phase everywhere, the faster choice at
 * 256^3 on MI355X; 1: pairs on levels large enough to fill the device (at
 * least 256 workgroups); 2: pairs on every level.  No reference
 * counterpart (the order of the phases is rbgsSmoother.C:376-437's either
 * way). */
int bmg_set_fused_smoothing(bmg_ctx* ctx, int mode);
/* Red-black smoothing on levels with periodic faces.  mode 1 (default): in
 * the FP32 cycles of mixed precision each phase covers the whole interior
 * and writes the periodic face ghosts together with the cells they mirror
 * (levels with an even count of at least 8 cells along every periodic
 * axis), so a sweep
 * is its phases only; the FP64 cycles keep the reference's order --
 * interior phases, ghost fill, then the boundary slabs
 * (rbgsSmoother.C:186-374) -- bit for bit.  mode 2: the wrapped phases in
 * the FP64 cycles as well (periodic Gauss-Seidel in another order than the
 * reference's: not bitwise equal to it).  mode 0: the reference's order in
 * the FP32 cycles too.  No reference counterpart. */
int bmg_set_periodic_smoothing(bmg_ctx* ctx, int mode);

#ifdef __cplusplus
}
#endif

#endif
