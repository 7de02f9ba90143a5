/*
 * bmg_oracle.h -- CPU restatement of Briscola's geometric-multigrid (MG)
 * pressure-Poisson solve, used ONLY as test infrastructure (the parity
 * checker and the reported CPU baseline).  Nothing in the product links or
 * calls this library.
 *
 * The restatement follows the reference files listed per function in
 * bmg_oracle.cpp.  Layout and conventions are the reference's own:
 *   - every level field is a ghost-padded block of (Nx+2)(Ny+2)(Nz+2)
 *     doubles, index lin(i,j,k) = (i*(Ny+2)+j)*(Nz+2)+k with padded indices
 *     (interior cell (0,0,0) sits at padded (1,1,1)), k (z) fastest
 *     (src/briscolaCore/include/utils.H:559-570);
 *   - the operator is the 7-component AoS stencil [C,L,R,B,T,A,F] per cell
 *     (src/briscolaCore/primitives/StencilSpace/stencil/stencil.H:21).
 */
#ifndef BMG_ORACLE_H
#define BMG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

/* Boundary kinds per face (-x,+x,-y,+y,-z,+z). */
enum { ORACLE_NEUMANN = 0, ORACLE_DIRICHLET = 1, ORACLE_PERIODIC = 2 };
/* MGCycleType order of src/briscolaFiniteVolume/solvers/MG/MG.H:18-23 */
enum { ORACLE_VCYCLE = 0, ORACLE_WCYCLE = 1, ORACLE_FCYCLE = 2 };
/* MGCoarseMode order of MG.H:26-30 */
enum { ORACLE_COARSE_SMOOTH = 0, ORACLE_COARSE_DIRECT = 1 };
enum { ORACLE_RBGS = 0, ORACLE_JAC = 1, ORACLE_LEXGS = 2 };
enum { ORACLE_X = 0, ORACLE_B = 1, ORACLE_R = 2 };

/* Build the level hierarchy and metrics from level-0 vertex coordinates
 * (n[d]+1 values per axis).  bc_value: Neumann gradient / Dirichlet value,
 * uniform per face, applied on level 0 only (per-cell data:
 * oracle_set_boundary_values). */
oracle_ctx* oracle_create(const int n[3], const double* vx, const double* vy,
                          const double* vz, const int bc[6],
                          const double bc_value[6]);
void oracle_destroy(oracle_ctx*);
void oracle_set_threads(int n);
int oracle_get_threads(void);

/* Solver dictionary (MG.C:257-304, solver.C:114-123). */
void oracle_set_opts(oracle_ctx*, double rel_tol, double tolerance,
                     int min_iter, int max_iter, int n_sweeps_pre,
                     int n_sweeps_post, int cycle, int coarse_mode,
                     int coarse_level, int smoother);

int oracle_num_levels(const oracle_ctx*);
void oracle_level_shape(const oracle_ctx*, int l, int n[3]);
void oracle_level_ratio(const oracle_ctx*, int l, int r[3]);
/* cell centres / widths per axis, padded index (N+2 values) */
void oracle_level_centres(const oracle_ctx*, int l, int axis, double* out);
void oracle_level_widths(const oracle_ctx*, int l, int axis, double* out);

/* imLaplacian on all levels (linearGaussLaplacianScheme.C:22-102).
 * lam: NULL for the lambda-free operator, else three padded level-0 arrays of
 * lower-face coefficients, restricted to coarse levels area-weighted. */
void oracle_assemble_laplacian(oracle_ctx*, const double* lam_x,
                               const double* lam_y, const double* lam_z);
/* The same minus im::ddt(x)/dt_frac of the Euler scheme: every interior
 * centre lowered by (((dx*dy)*dz)/delta_t)/dt_frac, every level
 * (defaultPoissonSolver.C:91-92, EulerDdtScheme.C:50-54).  delta_t = 0: no
 * time term. */
void oracle_assemble_laplacian_ddt(oracle_ctx*, const double* lam_x,
                                   const double* lam_y, const double* lam_z,
                                   double delta_t, double dt_frac);
void oracle_set_stencil(oracle_ctx*, int l, const double* A7);
void oracle_get_stencil(const oracle_ctx*, int l, double* A7);
void oracle_set_symmetric(oracle_ctx*, int symmetric);
int oracle_singular(const oracle_ctx*);

void oracle_set_field(oracle_ctx*, int which, int l, const double* v);
void oracle_get_field(const oracle_ctx*, int which, int l, double* v);

/* Pieces of MG::solve (MG.C:402-483) exposed for kernel-level parity. */
void oracle_pin_and_eliminate(oracle_ctx*);   /* MG.C:431-447 */
void oracle_prepare_coarse(oracle_ctx*);      /* APLU::prepare */
void oracle_smooth(oracle_ctx*, int l, int sweeps);
void oracle_residual(oracle_ctx*, int l, int corr_bcs);
void oracle_restrict(oracle_ctx*, int l);     /* b[l+1] = R r[l] (scaled) */
void oracle_prolong(oracle_ctx*, int l);      /* x[l] += P x[l+1] */
void oracle_correct_bcs(oracle_ctx*, int l);  /* correctBoundaryConditions */
void oracle_coarse_solve(oracle_ctx*);        /* APLU::solve */
void oracle_cycle(oracle_ctx*);               /* one MG::cycle from level 0 */
/* Per-level counts of the last cycle (oracle_cycle or the last iteration of
 * oracle_solve): what = 0 MG::cycle calls (visits), 1 Smooth calls,
 * 2 residuals, 3 restrictions, 4 prolongations, 5 coarse solves. */
void oracle_cycle_counts(const oracle_ctx*, int what, long* out);
/* Per-cell level-0 boundary data of a Neumann (gradient) or Dirichlet
 * (value) face: the patch block of the face's boundary layer, shape E - S
 * with the normal axis of size 1, index (ijk - S), last axis fastest
 * (boundaryValues_[d], DirichletBoundaryConditionBase.H:30;
 * boundaryGradients_[d], NeumannBoundaryConditionBase.H:30).  NULL: back to
 * the uniform value of oracle_create.  Returns 0, or -1 for a bad face. */
int oracle_set_boundary_values(oracle_ctx*, int face, const double* v);
void oracle_correct_eliminated(oracle_ctx*);  /* x[0].correct<eliminatedBcs> */
double oracle_sum_sq_residual(const oracle_ctx*);
void oracle_norm_factor(const oracle_ctx*, double* nf);

/* MG::solve of a diagonal system (MG.C:404-428). */
void oracle_solve_diagonal(oracle_ctx*);

/* Whole MG::solve: stats = {iters, initial, final}. */
void oracle_solve(oracle_ctx*, int const_matrix, double stats[3]);

/* Dense coarse system of the coarse level after pin and elimination, and its
 * APLU factors (row-major n*n, pivots n).  Returns n. */
int oracle_coarse_dense(const oracle_ctx*, double* dense, double* lu,
                        int* pivots);

#ifdef __cplusplus
}
#endif

#endif
