// Internal declarations shared by the HIP kernels and the host solver.
//
// Device layout ("colour-split"): every level field is stored as two arrays,
// one per red-black colour c = (I+J+K) & 1 of the padded cell index.  Cell
// (I,J,K) lives at position ((I*P1 + J)*Hz + (K>>1)) of its colour array,
// Hz = (P2+1)/2.  A red-black phase then streams only the cells of its
// colour, and the six neighbours of a cell are the same position in the
// rows (I+-1,J) and (I,J+-1) of the other colour plus positions
// p-1+q / p+q (q = K&1) of its own row there.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace bmg {

enum { NEU = 0, DIR = 1, PER = 2 };

// Operator representations:
//   OP_STENCIL: the reference AoS 7-point stencil, split by colour and stored
//               component-major (7 arrays per colour);
//   OP_CONST  : lambda-free rectilinear Laplacian, coefficients rebuilt from
//               1-D width/delta arrays on the fly (no per-cell bytes);
//   OP_FACES  : three lower-face coefficients per cell (area*delta*lambda).
enum { OP_STENCIL = 0, OP_CONST = 1, OP_FACES = 2 };

struct LevelDims {
    int N[3];
    int P[3];
    int Hz;
    int sI;        // colour-array stride of one I plane = P1*Hz
    long ncol;     // elements per colour array = P0*P1*Hz (< 2^31/8)
    int bc[6];
    int R[3];      // refinement ratio of this level vs its parent
};

// Storage of the stored face coefficients (OP_FACES): FP64 for the FP64
// path; FP16 scaled by a per-level factor for the FP32 cycles of mixed
// precision, whose smoothing only preconditions an FP64 iteration (half the
// bytes of the phase's largest streams).
template <typename T>
struct FaceStore { using type = T; };
template <>
struct FaceStore<float> { using type = _Float16; };

template <typename T>
struct OpView {
    int kind;
    const T* a7[2];   // OP_STENCIL: 7*ncol per colour, component-major
    // OP_FACES: 3*ncol per colour, component-major; FP32 coefficients are
    // the stored FP16 values times fs
    const typename FaceStore<T>::type* wf[2];
    T fs;
    const T* w[3];    // cell widths per padded index
    const T* dl[3];   // 1/|dcc| of the lower face of each padded index
    const T* dg[2];   // OP_CONST/OP_FACES: implicit Euler diagonal
                      // (V/dt)/dtFrac per colour, subtracted from the
                      // centre; null without a time term
};

template <typename T>
struct Pair {
    T* c[2];
};

// Restriction / prolongation per-axis data (in T).
template <typename T>
struct TransferView {
    const T* ru[3];   // coarse padded index -> u of the restriction weights
    const T* pu[3];   // fine padded index   -> u of the prolongation weights
    const T* fw[3];   // fine widths
    const T* cw[3];   // coarse widths
};

struct Region {
    int S[3];
    int E[3];
    // bit d set: the periodic face ghosts of axis d are written with their
    // cells (red-black phases of the FP32 cycles on levels with an even cell
    // count along every periodic axis); 0 everywhere else
    int wrap = 0;
    // red-black phases: 1 takes every neighbour as 0 without reading it (the
    // first phase from x = 0 of an FP32 cycle, whose other colour is never
    // read before the next phase rewrites it)
    int zeroNb = 0;
};

// Level-0 boundary data of the Neumann (gradient) and Dirichlet (value)
// faces, order -x,+x,-y,+y,-z,+z: p[f] null -> the uniform value v[f];
// else per boundary cell, device memory in the reference's patch-block
// layout (the face's boundary layer with the normal axis of size 1, index
// (ijk - S), last axis fastest): the value of interior cell (I,J,K) (padded
// indices) of the face normal to d is p[f][(t1 - 1)*N[a2] + (t2 - 1)] with
// (t1, t2) its indices along the two other axes a1 < a2.
// g[f], when set, holds the level-0 ghost coefficients of face f before
// elimination in the same layout (a stored stencil loses them when it is
// eliminated in place); otherwise they come from the operator.
struct BcData {
    double v[6];
    const double* p[6];
    const double* g[6];
};

// ---- batched execution (bmg_kernels.hip, bmg_batch.hip) ----
// One system's share of a batched launch: virtual workgroups
// [start, start + n) run the system's launch with grid (gx, gy, n/(gx*gy)).
// The argument struct of the share is args[arg].
struct BatchItem {
    unsigned start;
    unsigned gx, gy, n;
    int arg;
};
// The shares a batched launch walks, in device memory so that the host can
// drop systems between replays of a captured launch sequence.
struct BatchState {
    int nitems;
    unsigned total;           // virtual workgroups
};
// A batched launch of one kernel over many systems.
struct BatchLaunch {
    const void* args;         // device array of argument structs
    const BatchItem* items;   // device array, ascending starts
    const BatchState* state;
    unsigned grid;            // workgroups launched, a multiple of 8
    size_t shm;
};
struct LaunchRecord {
    const void* kind;
    long tag;          // position in the solve's structure (record_tag)
    unsigned grid[3];
    size_t shm;
    size_t align;
    std::vector<unsigned char> args;
    void (*many)(const BatchLaunch&, hipStream_t);
    // Rewrites the tile shape of the argument struct in place for rows of
    // 2^ws cells per wavefront row (3 <= ws <= 6) and sets its grid;
    // returns 0, changing nothing, when the kernel keeps its tiles (null:
    // the kernel has one tile shape).  The per-cell arithmetic is the same
    // under every shape.
    int (*retile)(unsigned char* args, int ws, unsigned grid[3]);
};
struct LaunchRecorder {
    std::vector<LaunchRecord> recs;
    long tag = 0;
};
// The calling thread's recorder: while it is set, the batchable launchers
// append records instead of launching and the others throw.
LaunchRecorder*& launch_recorder();
// Tags the launches recorded from now on with a position in the solve's
// structure (no effect without a recorder).  Launches of different systems
// with equal tags are merged with each other; tags must not decrease along
// one system's sequence for the tags to be used.
void record_tag(long tag);

// ---- launchers (bmg_kernels.hip) ----
// number of per-tile partial sums a cell-pair reduction over level L writes
int pair_tile_count(const LevelDims& L);
template <typename T>
void launch_rbgs(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                 Pair<const T> b, int colour, const Region& rg,
                 hipStream_t s);
// n (1..8) red-black phases over rg, colours cols[0..n), in one launch of
// one workgroup with a barrier between phases (the first phase with
// rg.zeroNb when zeroFirst): the same bits as n launch_rbgs calls.  For
// small levels; not batchable.
template <typename T>
void launch_rbgs_chain(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                       Pair<const T> b, const Region& rg, const int* cols,
                       int n, int zeroFirst, hipStream_t s);
// The phase and, with snap, the periodic ghost fill (mode 0) with its
// slab-plane snapshots in the same launch (the same bits as launch_rbgs
// followed by launch_bc_fill_snap).
template <typename T>
void launch_rbgs_fill(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                      Pair<const T> b, int colour, const Region& rg,
                      int level0, const BcData* bcval, T* const (*snap)[2],
                      hipStream_t s);
// Two consecutive red-black phases (colour c1, then 1 - c1) over rg in one
// launch: c1 updated in place in xR, the other colour read from xBin and
// written to xBout (a second buffer; cells outside rg copied).  Not
// batchable.  fused_rbgs_tiles: the launch's workgroup count, 0 when the
// level's rows are too long for its LDS ring.
template <typename T>
void launch_rbgs2(const LevelDims& L, const OpView<T>& op, T* xR,
                  const T* xBin, T* xBout, const T* bR, const T* bB, int c1,
                  const Region& rg, hipStream_t s);
template <typename T>
int fused_rbgs_tiles(const LevelDims& L);
// Slab-plane snapshots and the snapshots a slab launch writes on to: src
// holds the snapshot of each of its planes; every target is the snapshot of
// a later slab launch, of plane tplane normal to axis tdim.
template <typename T>
struct SlabIO {
    const T* src[2];
    int ntarget;
    T* tptr[6];
    int tdim[6];
    int tplane[6];
};
// All phases of the boundary slabs of one face pair (both planes when they
// are independent, nslab = 2) in one launch; colours[4*slab + phase].
template <typename T>
void launch_rbgs_slabs(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                       Pair<const T> b, int d, const int* planes, int nslab,
                       const int* colours, int nphase, const SlabIO<T>& io,
                       hipStream_t s);
template <typename T>
void launch_lexgs(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                  Pair<const T> b, const Region& rg, hipStream_t s);
template <typename T>
void launch_jacobi(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                   Pair<const T> b, Pair<T> tmp, const Region& rg,
                   hipStream_t s);
// r = b - A x over the interior (and per-tile partial sums of r^2 when
// partials is not null; *nblocks receives their number)
template <typename T>
void launch_residual(const LevelDims& L, const OpView<T>& op, Pair<const T> x,
                     Pair<const T> b, Pair<T> r, double* partials,
                     int* nblocks, hipStream_t s);
template <typename T>
void launch_restrict(const LevelDims& Lf, const LevelDims& Lc,
                     const TransferView<T>& tv, Pair<const T> rf, Pair<T> bc,
                     hipStream_t s);
template <typename T>
// skipColour (levels refined by 2 on every axis): the fine cells of that
// colour are left unchanged, for a caller whose next red-black phase
// overwrites every cell of it; -1 updates all
void launch_prolong(const LevelDims& Lf, const LevelDims& Lc,
                    const TransferView<T>& tv, Pair<const T> xc, Pair<T> xf,
                    hipStream_t s, int skipColour = -1);
// mode 0: periodic faces only (evaluate<nonEliminatedBcs,1>);
// mode 1: full correctBoundaryConditions; mode 2: Neumann/Dirichlet faces
// only (correct<eliminatedBcs>); level0 enables the inhomogeneous terms.
template <typename T>
void launch_bc_fill(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                    int mode, int level0, const BcData* bcval, hipStream_t s);
// The same, also copying the slab planes of every periodic axis d into
// snap[d][0] (plane 1) and snap[d][1] (plane N_d), P_a1 x P_a2 each
// (snap == nullptr or a null row: none for that axis).
template <typename T>
void launch_bc_fill_snap(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                         int mode, int level0, const BcData* bcval,
                         T* const (*snap)[2], hipStream_t s);
// Deflation by indicator vectors of cell sets (flexible CG): the entries
// (A z_k)_i of the level-0 operator (count only when cell is null), chunk
// partial sums part[c] = sum val * f(cell) over [beg[c], end[c]) of a list
// sorted by component (val null: plain sums of f), and the correction
// f(cell) += sign * (Einv w)[comp(cell)], w[k] the sum of the partials of
// chunks coff[k]..coff[k+1] (at most 64 components).
unsigned launch_defl_entries(const LevelDims& L, const OpView<double>& op,
                             const int* label, unsigned* count, unsigned cap,
                             unsigned* cell, int* comp, double* val,
                             hipStream_t s);
template <typename T>
void launch_defl_dot(const unsigned* cell, const double* val, const int* beg,
                     const int* end, int nchunks, Pair<const T> f,
                     double* part, hipStream_t s);
template <typename T>
void launch_defl_correct(const unsigned* cell, const int* comp, long ncells,
                         const double* einv, const double* part,
                         const int* coff, int m, double sign, Pair<T> f,
                         hipStream_t s);
// Coarsest-level solve by a precomputed inverse (column-major, 64 x 64,
// at most 64 cells), the ghost update after it.
template <typename T>
void launch_coarse_inv(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                       Pair<const T> b, const double* inv, int n, int level0,
                       const BcData* bcval, hipStream_t s);
// x = 0 everywhere, then the red-black phase of `colour` over rg from it,
// in one pass (the same bits as launch_fill followed by launch_rbgs).
template <typename T>
void launch_zero_phase(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                       Pair<const T> b, int colour, const Region& rg,
                       hipStream_t s);
template <typename T>
void launch_fill(Pair<T> x, long ncol, T v, hipStream_t s);
template <typename T>
void launch_ref_to_split(const LevelDims& L, const double* ref, Pair<T> out,
                         hipStream_t s);
template <typename T>
void launch_split_to_ref(const LevelDims& L, Pair<const T> in, double* ref,
                         hipStream_t s);
void launch_stencil_ref_to_split(const LevelDims& L, const double* ref7,
                                 Pair<double> a7, hipStream_t s);
void launch_stencil_split_to_ref(const LevelDims& L, Pair<const double> a7,
                                 double* ref7, hipStream_t s);
template <typename T>
void launch_materialize_stencil(const LevelDims& L, const OpView<T>& op,
                                Pair<T> a7, hipStream_t s);
void launch_pin(const LevelDims& L, Pair<double> a7, hipStream_t s);
void launch_eliminate(const LevelDims& L, Pair<double> a7, hipStream_t s);
// The ghost coefficient (component f+1) of every level-0 boundary cell of
// face f of a stored stencil, in the boundary-data layout of BcData.
void launch_save_ghost_coeffs(const LevelDims& L, Pair<const double> a7,
                              int f, double* out, hipStream_t s);
// b[0] -= inhomogeneous boundary terms of eliminateGhosts (level 0 only)
void launch_bc_source(const LevelDims& L, const OpView<double>& op,
                      Pair<double> b, const BcData* bcval, hipStream_t s);
void launch_lambda_restrict(const LevelDims& Lf, const LevelDims& Lc,
                            const double* fw0, const double* fw1,
                            const double* fw2, const double* lamf,
                            double* lamc, int fd, hipStream_t s);
void launch_split_source(const LevelDims& L, const double* x,
                         const double* const lam[3], double lam0,
                         const double* b, const double* const w[3],
                         const double* const dl[3], double* const phi[3],
                         double* bhat, hipStream_t s);
void launch_ddt_diagonal(const LevelDims& L, const double* w0,
                         const double* w1, const double* w2, double delta_t,
                         double dt_frac, Pair<double> dg, hipStream_t s);
void launch_face_coeffs(const LevelDims& L, const double* w0,
                        const double* w1, const double* w2, const double* dl,
                        const double* lam, int fd, Pair<double> wf,
                        hipStream_t s);
// Flux through the lower faces normal to fd on level-0 arrays in the
// reference layout: flux = (lam*((x_lower - x)*dl))*A (lam may be null).
void launch_face_flux(const LevelDims& L, const double* x, const double* lam,
                      const double* const w[3], const double* dl, int fd,
                      double* flux, hipStream_t s);
// Width of the one-wavefront coarse solve for an n-cell coarse level (a
// power of two >= n, at least 8), or 0 when n exceeds one wavefront.  With a
// width N the factors are stored column-major in an N x N array padded with
// unit rows (lu[j*N + i] = LU(i, j)) and piv holds the row permutation
// (row i of the permuted right-hand side is row piv[i] of the original);
// otherwise lu is the n x n row-major LU and piv the pivot sequence.
inline int coarse_wave_width(int n)
{
    if (n > 64) return 0;
    int w = 8;
    while (w < n) w *= 2;
    return w;
}
template <typename T>
void launch_coarse_solve(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                         Pair<const T> b, const double* lu, const int* piv,
                         int n, int level0, const BcData* bcval,
                         hipStream_t s);
template <typename T>
void launch_sum_sq(const LevelDims& L, Pair<const T> f, double* partials,
                   int* nblocks, hipStream_t s);
template <typename T>
void launch_sum(const LevelDims& L, Pair<const T> f, double* partials,
                int* nblocks, hipStream_t s);
template <typename T>
void launch_norm_factor_terms(const LevelDims& L, const OpView<T>& op,
                              Pair<const T> r, Pair<const T> b, double y,
                              double* partials, int* nblocks, hipStream_t s);
// out[0] = sum of n partials; also written to hostOut (host-mapped pinned
// memory, read by the host after the stream synchronises) when given
void launch_reduce_final(const double* partials, int n, double* out,
                         hipStream_t s, double* hostOut = nullptr);
void launch_diag(const LevelDims& L, Pair<const double> a7, Pair<double> x,
                 Pair<const double> b, hipStream_t s);
// Flexible CG with device-resident scalars sc[0..3] = p.q, p.r, r.r, z.q.
// q = A p with the partials of p.q (row 0) and p.r (row 1); returns tiles.
int launch_apply_dots(const LevelDims& L, const OpView<double>& op,
                      Pair<const double> p, Pair<double> q,
                      Pair<const double> r, double* partials, hipStream_t s);
// r -= alpha q (alpha = sc[1]/sc[0]), partials of r.r, and r copied for the
// preconditioner: to FP32 rf when rf.c[0] is set, else to the FP64 r0 and b0
// when r0.c[0] is set, else nowhere.  Returns tiles.  (x += alpha p follows
// in launch_update_p, which reads p anyway.)
int launch_fcg_update(const LevelDims& L, const double* sc, Pair<double> r,
                      Pair<const double> q, Pair<float> rf, Pair<double> r0,
                      Pair<double> b0, double* partials, hipStream_t s);
template <typename Z>
int launch_dot_zq(const LevelDims& L, Pair<const Z> z, Pair<const double> q,
                  double* partials, hipStream_t s);
// x += alpha p and then p = z + beta p on interior cells, alpha =
// sc[1]/sc[0] (the step of the direction p just used), beta = -sc[3]/sc[0]
template <typename Z>
void launch_update_p(const LevelDims& L, const double* sc, Pair<const Z> z,
                     Pair<double> p, Pair<double> x, hipStream_t s);
// level-0 interior <-> dense FP32 (reference cell order) for the learned
// preconditioner: out = r / rms(r), z = in * rms / h, rms = sqrt(ss/cells)
// with ss = *ssDev when ssDev is set, else ssHost
void launch_split_to_dense(const LevelDims& L, Pair<const double> r,
                           const double* ssDev, double ssHost, float* out,
                           hipStream_t s);
void launch_dense_to_split(const LevelDims& L, const float* in,
                           const double* ssDev, double ssHost, double invh,
                           Pair<double> z, hipStream_t s);
// x += z on interior cells
void launch_add_interior(const LevelDims& L, Pair<double> x,
                         Pair<const double> z, hipStream_t s);
// out[row] = sum of partials[row*n .. row*n + n) for row < nrows
void launch_reduce_rows(const double* partials, int n, int nrows, double* out,
                        hipStream_t s, double* hostOut = nullptr);
template <typename T>
void launch_copy(long n, Pair<const T> a, Pair<T> b, hipStream_t s);
void launch_to_double(long n, Pair<const float> a, Pair<double> b,
                      hipStream_t s);
// x64 += e32 (interior), r32 = (float) r64, etc.
void launch_add_correction(const LevelDims& L, Pair<double> x,
                           Pair<const float> e, hipStream_t s);
// Largest |value| of the first n elements of both arrays of a pair.
double device_absmax(long n, Pair<const double> in, double* partials,
                     size_t npartials, hipStream_t s);
// out = (FP16) (in * scale), n elements per colour.
void launch_to_half(long n, Pair<const double> in, double scale,
                    Pair<_Float16> out, hipStream_t s);
void launch_to_float(long ncol, Pair<const double> in, Pair<float> out,
                     hipStream_t s);
template <typename T>
void launch_set_cell(Pair<T> f, long idx, int colour, T v, hipStream_t s);

}  // namespace bmg
