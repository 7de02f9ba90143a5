// HIP kernels of the MI355X multigrid solver (gfx950).
//
// All kernels work on the colour-split layout described in bmg_internal.h.
// They are built with -ffp-contract=off: every product and sum is rounded in
// the order the reference writes it, so the FP64 path reproduces the
// reference arithmetic bit for bit (the order of each expression is cited at
// the kernel).
#include "bmg_internal.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace bmg {

namespace {

constexpr int BX = 64;   // one wavefront along the contiguous axis
constexpr int BY = 4;

using BcVals = BcData;

// Boundary value of face f at interior cell s (padded indices) of the
// face's boundary layer: the per-cell array when set (DirichletBoundary-
// Condition.C:45,68, NeumannBoundaryCondition.C:51,83 read
// boundaryValues_[d](ijk - S)), else the face's uniform value.
__device__ __forceinline__ double bc_value_at(const LevelDims& L,
                                              const double* p, int f,
                                              const int s[3])
{
    const int d = f >> 1;
    const int a1 = d == 0 ? 1 : 0, a2 = d == 2 ? 1 : 2;
    return p[(s[a1] - 1)*L.N[a2] + (s[a2] - 1)];
}

__device__ __forceinline__ double bc_value(const LevelDims& L,
                                           const BcVals& bv, int f,
                                           const int s[3])
{
    return bv.p[f] ? bc_value_at(L, bv.p[f], f, s) : bv.v[f];
}


__device__ __forceinline__ int cidx(const LevelDims& L, int I, int J, int K)
{
    return (I*L.P[1] + J)*L.Hz + (K >> 1);
}

__device__ __forceinline__ int nci(const LevelDims& L) { return int(L.ncol); }

// Ghost coefficient of every boundary cell of face f of a stored level-0
// stencil (component f+1, before elimination), one thread per cell.
__global__ void k_save_ghost_coeffs(LevelDims L, Pair<const double> a7,
                                    int f, double* out)
{
    const int d = f >> 1;
    const int a1 = d == 0 ? 1 : 0, a2 = d == 2 ? 1 : 2;
    const int n1 = L.N[a1], n2 = L.N[a2];
    const int t = int(blockIdx.x*blockDim.x + threadIdx.x);
    if (t >= n1*n2) return;
    int s[3];
    s[a1] = 1 + t/n2;
    s[a2] = 1 + t%n2;
    s[d] = (f & 1) ? L.N[d] : 1;
    const int c = (s[0] + s[1] + s[2]) & 1;
    out[t] = (c ? a7.c[1] : a7.c[0])[size_t(f + 1)*size_t(L.ncol) +
                                     size_t(cidx(L, s[0], s[1], s[2]))];
}

// Element i of an array addressed by a 32-bit byte offset from a base held
// in scalar registers (every array is below 2^31 bytes), so that a load is
// one global_load with a 32-bit VGPR offset instead of a 64-bit address add.
template <typename T>
__device__ __forceinline__ T ld(const T* base, unsigned i)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                       size_t(i*unsigned(sizeof(T))));
}

// Array of colour c of a field.  A select between the two kernel-argument
// pointers (rather than an index into the argument) keeps the result a
// global-memory pointer.
template <typename T>
__device__ __forceinline__ T* col(const Pair<T>& x, int c)
{
    return c ? x.c[1] : x.c[0];
}

template <typename T>
__device__ __forceinline__ void st(T* base, unsigned i, T v)
{
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) +
                          size_t(i*unsigned(sizeof(T)))) = v;
}

// Stencil coefficients [C,L,R,B,T,A,F] of padded cell (I,J,K), as assembled
// by linearGaussLaplacianScheme.C:61-94 (centre = 0 - v_xl - v_xu - v_yl -
// v_yu - v_zl - v_zu, faces in loop order), before ghost elimination.
// withDdt false leaves out the time term's diagonal (ddt_centre).
template <typename T, int OP>
__device__ __forceinline__ void raw_coeffs(const LevelDims& L,
                                           const OpView<T>& op, int colour,
                                           int idx, int row, int p, int q,
                                           int I, int J, int K, T a[7],
                                           bool withDdt = true)
{
    if constexpr (OP == OP_STENCIL)
    {
        const T* s = op.a7[colour];
        const int n = nci(L);
#pragma unroll
        for (int c = 0; c < 7; c++) a[c] = ld(s, unsigned(c*n + idx));
    }
    else
    {
        if constexpr (OP == OP_CONST)
        {
            const T wx = ld(op.w[0], I), wy = ld(op.w[1], J);
            const T wz = ld(op.w[2], K);
            const T ax = wy*wz;
            const T ay = wx*wz;
            const T az = wx*wy;
            a[1] = ax*ld(op.dl[0], I);
            a[2] = ax*ld(op.dl[0], I + 1);
            a[3] = ay*ld(op.dl[1], J);
            a[4] = ay*ld(op.dl[1], J + 1);
            a[5] = az*ld(op.dl[2], K);
            a[6] = az*ld(op.dl[2], K + 1);
        }
        else
        {
            const int n = nci(L);
            using F = typename FaceStore<T>::type;
            const F* wc = op.wf[colour];
            const F* wo = op.wf[colour ^ 1];
            const F f[6] = {ld(wc, idx), ld(wo, idx + L.sI), ld(wc, n + idx),
                            ld(wo, n + idx + L.Hz), ld(wc, 2*n + idx),
                            ld(wo, 2*n + row + p + q)};
#pragma unroll
            for (int s = 0; s < 6; s++)
            {
                if constexpr (std::is_same<F, T>::value)
                    a[s + 1] = f[s];
                else
                    a[s + 1] = T(f[s])*op.fs;
            }
        }
        T c0 = T(0) - a[1];
        c0 = c0 - a[2];
        c0 = c0 - a[3];
        c0 = c0 - a[4];
        c0 = c0 - a[5];
        c0 = c0 - a[6];
        // sys -= ddt(x)/dtFrac lowers the centre by the Euler diagonal
        // (defaultPoissonSolver.C:91-92, EulerDdtScheme.C:53)
        if (withDdt && op.dg[0]) c0 = c0 - ld(op.dg[colour], idx);
        a[0] = c0;
    }
}

// Neumann/Dirichlet eliminateGhosts for one cell, faces in order 0..5
// (NeumannBoundaryCondition.C:40-48: C += g; DirichletBoundaryCondition.C:
// 36-43: C -= g; the ghost coefficient becomes 0).
template <typename T>
__device__ __forceinline__ void eliminate_cell(const LevelDims& L, int I,
                                               int J, int K, T a[7])
{
    const bool on[6] = {I == 1, I == L.N[0], J == 1, J == L.N[1],
                        K == 1, K == L.N[2]};
    // most cells touch no eliminated face: branch around the whole block
    const bool faceI = (on[0] && L.bc[0] != PER) || (on[1] && L.bc[1] != PER);
    const bool faceJ = (on[2] && L.bc[2] != PER) || (on[3] && L.bc[3] != PER);
    const bool faceK = (on[4] && L.bc[4] != PER) || (on[5] && L.bc[5] != PER);
    if (!(faceI || faceJ || faceK)) return;
#pragma unroll
    for (int f = 0; f < 6; f++)
    {
        if (on[f] && L.bc[f] != PER)
        {
            const T g = a[f + 1];
            a[0] = (L.bc[f] == NEU) ? a[0] + g : a[0] - g;
            a[f + 1] = T(0);
        }
    }
}

template <typename T, int OP>
__device__ __forceinline__ void coeffs(const LevelDims& L, const OpView<T>& op,
                                       int colour, int idx, int row, int p,
                                       int q, int I, int J, int K, T a[7])
{
    raw_coeffs<T, OP>(L, op, colour, idx, row, p, q, I, J, K, a);
    if constexpr (OP != OP_STENCIL)
        eliminate_cell(L, I, J, K, a);
}

__device__ __forceinline__ double block_sum(double v, double* sm)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    const int tid = threadIdx.x + threadIdx.y*blockDim.x;
    const int nt = blockDim.x*blockDim.y;
    if ((tid & 63) == 0) sm[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
        for (int w = 0; w < (nt + 63)/64; w++) s += sm[w];
    return s;
}

// Workgroup index of a kernel body: blockIdx for a single launch, the
// system-local index in a batched launch.
struct Bid {
    unsigned x, y, z;
};

// A logical 3-D grid of tiles launched as a 1-D grid.  The device hands
// workgroups to its eight XCDs in turn (workgroup w to XCD w mod 8); tile
// (w mod 8)*per + w/8 gives every XCD one contiguous run of tiles, so the
// planes a stencil reads again for the next plane are still in the L2 of
// the XCD that reads them.  Tile order is x fastest, then y, then z.
struct Tiles {
    int gx, gy, gz;
    int per;   // tiles per XCD
    int ws;    // log2 of the cells of a row a wavefront covers (6: one row
               // segment of 64 per wavefront; below 6, 64 >> ws rows)
    int rpt = 1;   // red-black phases: rows per thread (1, 2 or 4; a tile
                   // covers rpt times the rows, thread rows BY wavefront
                   // rows apart)
};

__device__ __forceinline__ bool tile_of(const Tiles& t, unsigned w, int& bx,
                                        int& by, int& bz)
{
    const unsigned lin = (w & 7u)*unsigned(t.per) + (w >> 3);
    if (lin >= unsigned(t.gx*t.gy*t.gz)) return false;
    const unsigned r = lin/unsigned(t.gx);
    bx = int(lin - r*unsigned(t.gx));
    bz = int(r/unsigned(t.gy));
    by = int(r - unsigned(bz)*unsigned(t.gy));
    return true;
}

__device__ __forceinline__ int tile_index(const Tiles& t, int bx, int by,
                                          int bz)
{
    return bx + t.gx*(by + t.gy*bz);
}

// Tiles over nplanes planes of nrows rows of nper positions, a wavefront
// covering 64 >> ws rows of 2^ws positions (ws = 6: one row segment of 64).
Tiles packed_tiles(int nper, int nrows, int nplanes, int ws);

// Cell-pair kernels: thread (p, J) owns the two cells K = 2p+1 and
// K = 2p+2 of padded row (I,J).  Cell A (K = 2p+1) sits at position p of
// colour cA = (I+J+1)&1, cell B (K = 2p+2) at position p+1 of colour
// cB = (I+J)&1; p runs over (N2+1)/2 pairs, so every lane of a full row
// holds two interior cells.  Tiles are BX pairs x BY rows, one tile row per
// plane I: a wavefront is 64 pairs of one row, so J, cA and cB are uniform.
struct PairPos {
    int I, J, p;
    int cA, cB;
    int row;          // first position of row (I,J)
    unsigned idx;     // row + p: cell A at idx of colour cA, cell B at idx+1
                      // of colour cB
    bool vB;          // K = 2p+2 is an interior cell
};

// PACK: a wavefront covers 64 >> ws rows of 2^ws pairs (tiles of
// pair_tiles_packed); else one row segment of 64 pairs, J uniform.
template <bool PACK = false>
__device__ __forceinline__ bool pair_pos(const LevelDims& L, int bx, int by,
                                         int bz, PairPos& c, int ws = 6)
{
    c.I = 1 + bz;
    if (PACK)
    {
        const int lane = int(threadIdx.x);
        c.J = 1 + (by*BY + int(threadIdx.y))*(64 >> ws) + (lane >> ws);
        c.p = (bx << ws) + (lane & ((1 << ws) - 1));
    }
    else
    {
        c.J = __builtin_amdgcn_readfirstlane(1 + by*BY + int(threadIdx.y));
        c.p = bx*BX + int(threadIdx.x);
    }
    if (c.J > L.N[1] || c.p >= (L.N[2] + 1)/2) return false;
    c.cB = (c.I + c.J) & 1;
    c.cA = c.cB ^ 1;
    c.row = (c.I*L.P[1] + c.J)*L.Hz;
    c.idx = unsigned(c.row + c.p);
    c.vB = 2*c.p + 2 <= L.N[2];
    return true;
}

// A x for both cells of a pair (row product in the C,L,R,B,T,A,F order,
// linearSystemFunctions.H:42-49).  All twelve x values are loaded before any
// arithmetic.
template <typename T, int OP>
__device__ __forceinline__ void pair_Ax(const LevelDims& L,
                                        const OpView<T>& op, Pair<const T> x,
                                        const PairPos& c, T Ax[2])
{
    const unsigned sI = L.sI, Hz = L.Hz, i = c.idx;
    // uniform selects between the two kernel-argument pointers keep the
    // loads in the global address space
    const T* xa = col(x, c.cA);
    const T* xb = col(x, c.cB);
    // cell A (K = 2p+1): own xa[i]; neighbours in xb: -i,+i,-j,+j at i,
    // -k = K 2p at xb[i], +k = K 2p+2 at xb[i+1] (cell B)
    const T uA = ld(xa, i), uB = ld(xb, i + 1);
    const T nA[6] = {ld(xb, i - sI), ld(xb, i + sI), ld(xb, i - Hz),
                     ld(xb, i + Hz), ld(xb, i), uB};
    // cell B (K = 2p+2): own xb[i+1]; neighbours in xa at i+1, -k = cell A,
    // +k = K 2p+3 at xa[i+1]
    const T nB[6] = {ld(xa, i + 1 - sI), ld(xa, i + 1 + sI),
                     ld(xa, i + 1 - Hz), ld(xa, i + 1 + Hz), uA,
                     ld(xa, i + 1)};
    // the coefficients of both cells, B's also where the pair has no cell B
    // (odd N along z: B is then the ghost K = N+1, whose entries and those of
    // its neighbours lie inside the padded arrays) and its product is
    // dropped; the time term's diagonals of both are loaded under one
    // branch after every other load, so no load waits for a branch
    T a[7], b[7];
    raw_coeffs<T, OP>(L, op, c.cA, int(i), c.row, c.p, 1, c.I, c.J, 2*c.p + 1,
                      a, false);
    raw_coeffs<T, OP>(L, op, c.cB, int(i) + 1, c.row, c.p + 1, 0, c.I, c.J,
                      2*c.p + 2, b, false);
    if constexpr (OP != OP_STENCIL)
    {
        if (op.dg[0])
        {
            const T dA = ld(op.dg[c.cA], i), dB = ld(op.dg[c.cB], i + 1);
            a[0] = a[0] - dA;
            b[0] = b[0] - dB;
        }
        eliminate_cell(L, c.I, c.J, 2*c.p + 1, a);
        eliminate_cell(L, c.I, c.J, 2*c.p + 2, b);
    }
    Ax[0] = a[0]*uA + a[1]*nA[0] + a[2]*nA[1] + a[3]*nA[2] + a[4]*nA[3]
          + a[5]*nA[4] + a[6]*nA[5];
    const T axB = b[0]*uB + b[1]*nB[0] + b[2]*nB[1] + b[3]*nB[2] + b[4]*nB[3]
                + b[5]*nB[4] + b[6]*nB[5];
    Ax[1] = c.vB ? axB : T(0);
}

// The Gauss-Seidel update x = (b - P)/A0, P = A1 x_-i + A2 x_+i + A3 x_-j
// + A4 x_+j + A5 x_-k + A6 x_+k (rbgsSmoother.C:65-183).  In FP64 every
// product and sum is rounded in that order and the division is correctly
// rounded: the reference's arithmetic, bit for bit.  The FP32 cycles of
// mixed precision carry no bitwise contract (their result only corrects an
// FP64 iterate): the row product is contracted into fused multiply-adds
// and the division is a multiplication by the hardware reciprocal.
template <typename T>
__device__ __forceinline__ T gs_update(const T a[7], const T nb[6], T bv)
{
    const T P =
        a[1]*nb[0]
      + a[2]*nb[1]
      + a[3]*nb[2]
      + a[4]*nb[3]
      + a[5]*nb[4]
      + a[6]*nb[5];
    return (bv - P)/a[0];
}

template <>
__device__ __forceinline__ float gs_update<float>(const float a[7],
                                                  const float nb[6], float bv)
{
#pragma clang fp contract(fast)
    const float P =
        a[1]*nb[0]
      + a[2]*nb[1]
      + a[3]*nb[2]
      + a[4]*nb[3]
      + a[5]*nb[4]
      + a[6]*nb[5];
    return (bv - P)*__builtin_amdgcn_rcpf(a[0]);
}

// Copies of a cell's new value into the periodic face ghosts that mirror
// it, for the axes set in wrap.  With an even cell count N along a periodic
// axis the ghost at 0 mirrors cell N and the ghost at N+1 cell 1, both of
// the cell's colour, so a phase of one colour keeps every ghost its
// neighbours read up to date and no ghost fill or boundary-slab pass is
// needed between phases.
template <typename T>
__device__ __forceinline__ void wrap_ghosts(const LevelDims& L,
                                            T* __restrict__ xc, unsigned idx,
                                            int I, int J, int K, int wrap,
                                            T v)
{
    if ((wrap & 1) && (I == 1 || I == L.N[0]))
    {
        const unsigned d = unsigned(L.N[0])*L.sI;
        st(xc, I == 1 ? idx + d : idx - d, v);
    }
    if ((wrap & 2) && (J == 1 || J == L.N[1]))
    {
        const unsigned d = unsigned(L.N[1])*unsigned(L.Hz);
        st(xc, J == 1 ? idx + d : idx - d, v);
    }
    if ((wrap & 4) && (K == 1 || K == L.N[2]))
    {
        const unsigned d = unsigned(L.N[2] >> 1);
        st(xc, K == 1 ? idx + d : idx - d, v);
    }
}

// One cell of a red-black phase: row J of plane I, the off-th cell of the
// colour in the row.
template <typename T, int OP>
__device__ __forceinline__ void rbgs_cell(const LevelDims& L,
                                          const OpView<T>& op,
                                          T* __restrict__ xc,
                                          const T* __restrict__ xo,
                                          const T* __restrict__ bcl,
                                          int colour, const Region& rg, int I,
                                          int J, int off)
{
    if (J >= rg.E[1]) return;
    const int q = (colour - I - J) & 1;
    const int pmin = (rg.S[2] - q + 1) >> 1;
    const int p = pmin + off;
    const int K = 2*p + q;
    if (K >= rg.E[2]) return;
    const unsigned sI = L.sI, Hz = L.Hz;
    const int row = (I*L.P[1] + J)*L.Hz;
    const unsigned idx = unsigned(row + p);
    // issue every streamed load before the coefficient arithmetic
    T nb[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    if (!rg.zeroNb)
    {
        nb[0] = ld(xo, idx - sI);
        nb[1] = ld(xo, idx + sI);
        nb[2] = ld(xo, idx - Hz);
        nb[3] = ld(xo, idx + Hz);
        nb[4] = ld(xo, idx - 1 + q);
        nb[5] = ld(xo, idx + q);
    }
    const T bv = ld(bcl, idx);
    T a[7];
    coeffs<T, OP>(L, op, colour, int(idx), row, p, q, I, J, K, a);
    const T v = gs_update<T>(a, nb, bv);
    st(xc, idx, v);
    if (rg.wrap) wrap_ghosts(L, xc, idx, I, J, K, rg.wrap, v);
}

// R cells of a red-black phase, rows J[0..R) of plane I at offsets off[]
// of the colour: every load of the R cells is issued before the arithmetic
// of any, R times the bytes in flight per wavefront.  The per-cell
// arithmetic is rbgs_cell's.
template <typename T, int OP, int R>
__device__ __forceinline__ void rbgs_cells(const LevelDims& L,
                                           const OpView<T>& op,
                                           T* __restrict__ xc,
                                           const T* __restrict__ xo,
                                           const T* __restrict__ bcl,
                                           int colour, const Region& rg,
                                           int I, const int Jin[R],
                                           const int off[R])
{
    int J[R], p[R], q[R], K[R], row[R];
    bool ok[R];
    int src = -1;
#pragma unroll
    for (int c = 0; c < R; c++)
    {
        J[c] = Jin[c];
        q[c] = (colour - I - J[c]) & 1;
        p[c] = ((rg.S[2] - q[c] + 1) >> 1) + off[c];
        K[c] = 2*p[c] + q[c];
        ok[c] = J[c] < rg.E[1] && K[c] < rg.E[2];
        if (ok[c] && src < 0) src = c;
    }
    if (src < 0) return;
    // a cell outside the region takes a valid one's position: its loads stay
    // in bounds and its result is not stored
#pragma unroll
    for (int c = 0; c < R; c++)
        if (!ok[c])
        {
            J[c] = J[src];
            q[c] = q[src];
            p[c] = p[src];
            K[c] = K[src];
        }
    const unsigned sI = L.sI, Hz = L.Hz;
    unsigned idx[R];
    T nb[R][6], bv[R];
#pragma unroll
    for (int c = 0; c < R; c++)
    {
        row[c] = (I*L.P[1] + J[c])*L.Hz;
        idx[c] = unsigned(row[c] + p[c]);
        const unsigned i = idx[c];
        if (rg.zeroNb)
        {
#pragma unroll
            for (int e = 0; e < 6; e++) nb[c][e] = T(0);
        }
        else
        {
            nb[c][0] = ld(xo, i - sI);
            nb[c][1] = ld(xo, i + sI);
            nb[c][2] = ld(xo, i - Hz);
            nb[c][3] = ld(xo, i + Hz);
            nb[c][4] = ld(xo, i - 1 + q[c]);
            nb[c][5] = ld(xo, i + q[c]);
        }
        bv[c] = ld(bcl, i);
    }
    T a[R][7];
#pragma unroll
    for (int c = 0; c < R; c++)
        coeffs<T, OP>(L, op, colour, int(idx[c]), row[c], p[c], q[c], I, J[c],
                      K[c], a[c]);
#pragma unroll
    for (int c = 0; c < R; c++)
    {
        const T v = gs_update<T>(a[c], nb[c], bv[c]);
        if (ok[c])
        {
            st(xc, idx[c], v);
            if (rg.wrap) wrap_ghosts(L, xc, idx[c], I, J[c], K[c], rg.wrap, v);
        }
    }
}

// ---------------------------------------------------------------------------
// Red-black Gauss-Seidel phase: update every cell of one colour inside
// region [S,E) (rbgsSmoother.C:65-183):
//   x_c = (b_c - (A1 x_-i + A2 x_+i + A3 x_-j + A4 x_+j + A5 x_-k + A6 x_+k))
//         / A0
// One thread per cell of the colour; lanes run along the contiguous axis;
// tiles of BX positions x BY rows, one tile row per plane, dealt to the
// XCDs in contiguous runs of planes.  Rows of at most 32 cells of a colour
// are packed: a wavefront covers 64 >> ws rows of 2^ws cells (tl.ws < 6),
// so short rows do not leave most lanes of a wavefront idle.
template <typename T, int OP>
__device__ __forceinline__ void rbgs_body(const Bid bid, const LevelDims& L,
                                          const OpView<T>& op,
                                          T* __restrict__ xc,
                                          const T* __restrict__ xo,
                                          const T* __restrict__ bcl,
                                          int colour, const Region& rg,
                                          const Tiles& tl)
{
    int bx, by, bz;
    if (!tile_of(tl, bid.x, bx, by, bz)) return;
    const int I = rg.S[0] + bz;
    // several rows per thread in FP32 only (rbgs_tiles): the FP64 kernels
    // are compiled without that path, which would set their register count
    // (and so their occupancy) for every launch
    if (std::is_same<T, float>::value && tl.rpt > 1)
    {
        // rpt rows per thread, BY wavefront rows apart
        const int ws = tl.ws;
        const int lane = int(threadIdx.x);
        const int wr = 64 >> ws;   // rows per wavefront row
        int J0, off0;
        if (ws == 6)
        {
            J0 = __builtin_amdgcn_readfirstlane(rg.S[1] + tl.rpt*by*BY +
                                                int(threadIdx.y));
            off0 = bx*BX + lane;
        }
        else
        {
            J0 = rg.S[1] + (tl.rpt*by*BY + int(threadIdx.y))*wr + (lane >> ws);
            off0 = (bx << ws) + (lane & ((1 << ws) - 1));
        }
        if (tl.rpt == 2)
        {
            const int J[2] = {J0, J0 + BY*wr}, off[2] = {off0, off0};
            rbgs_cells<T, OP, 2>(L, op, xc, xo, bcl, colour, rg, I, J, off);
        }
        else
        {
            const int J[4] = {J0, J0 + BY*wr, J0 + 2*BY*wr, J0 + 3*BY*wr};
            const int off[4] = {off0, off0, off0, off0};
            rbgs_cells<T, OP, 4>(L, op, xc, xo, bcl, colour, rg, I, J, off);
        }
        return;
    }
    if (tl.ws == 6)
    {
        // a wavefront is one row (BX = 64), so J and everything derived
        // from it is uniform and lives in scalar registers
        const int J = __builtin_amdgcn_readfirstlane(rg.S[1] + by*BY +
                                                     int(threadIdx.y));
        rbgs_cell<T, OP>(L, op, xc, xo, bcl, colour, rg, I, J,
                         bx*BX + int(threadIdx.x));
    }
    else
    {
        const int ws = tl.ws;
        const int lane = int(threadIdx.x);
        const int J = rg.S[1] + (by*BY + int(threadIdx.y))*(64 >> ws) +
                      (lane >> ws);
        rbgs_cell<T, OP>(L, op, xc, xo, bcl, colour, rg, I, J,
                         (bx << ws) + (lane & ((1 << ws) - 1)));
    }
}

template <typename T, int OP>
__global__ __launch_bounds__(BX*BY) void k_rbgs(LevelDims L, OpView<T> op,
                                                T* __restrict__ xc,
                                                const T* __restrict__ xo,
                                                const T* __restrict__ bcl,
                                                int colour, Region rg,
                                                Tiles tl)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    rbgs_body<T, OP>(bid, L, op, xc, xo, bcl, colour, rg, tl);
}

template <typename T, int OP>
struct KRbgs
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims L;
        OpView<T> op;
        T* xc;
        const T* xo;
        const T* bcl;
        int colour;
        Region rg;
        Tiles tl;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        rbgs_body<T, OP>(bid, a.L, a.op, a.xc, a.xo, a.bcl, a.colour, a.rg,
                         a.tl);
    }
    static bool retile(Args& a, int ws)
    {
        const int nK = a.rg.E[2] - a.rg.S[2];
        a.tl = packed_tiles((nK + 1)/2, a.rg.E[1] - a.rg.S[1],
                            a.rg.E[0] - a.rg.S[0], ws);
        return true;
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_rbgs<T, OP>), g, dim3(bx, by), shm, s, a.L, a.op,
                           a.xc, a.xo, a.bcl, a.colour, a.rg, a.tl);
    }
};


// Several red-black phases of one small level in a single workgroup of
// 1024 threads: for each phase the threads walk the phase's cells (plane,
// row, position among the row's cells of the colour) with the per-cell
// update of k_rbgs (rbgs_cell), then wait at a barrier, after which the next
// phase's reads of the other colour see every new value.  On levels of at
// most 16^3 cells a launch costs more than its phase's work (about 3.5 us
// per launch in a captured graph against 0.6 us per phase at 8^3 and 2.2 us
// at 16^3 here, measured).
struct RbgsChain
{
    int n;           // phases, at most 8
    int col[8];      // colour of each phase
    int zeroFirst;   // the first phase takes its neighbours as 0
};

template <typename T, int OP>
__global__ __launch_bounds__(1024) void k_rbgs_chain(LevelDims L,
                                                     OpView<T> op, Pair<T> x,
                                                     Pair<const T> b,
                                                     Region rg, RbgsChain ch)
{
    const int nI = rg.E[0] - rg.S[0], nJ = rg.E[1] - rg.S[1];
    const int nOff = (rg.E[2] - rg.S[2] + 1)/2;
    const int total = nI*nJ*nOff;
    for (int k = 0; k < ch.n; k++)
    {
        const int c = ch.col[k];
        Region r = rg;
        r.zeroNb = (k == 0 && ch.zeroFirst) ? 1 : 0;
        for (int t = int(threadIdx.x); t < total; t += int(blockDim.x))
        {
            const int off = t % nOff;
            const int rr = t/nOff;
            const int J = rg.S[1] + rr % nJ;
            const int I = rg.S[0] + rr/nJ;
            rbgs_cell<T, OP>(L, op, x.c[c], x.c[c ^ 1], b.c[c], c, r, I, J,
                             off);
        }
        __syncthreads();
    }
}


// Two consecutive red-black phases in one launch: colour c1 over region rg,
// then colour c2 = 1 - c1 over rg (rbgsSmoother.C:65-183), with the per-cell
// arithmetic of rbgs_body.  Workgroup (jt, it) owns padded rows
// [j0, j0 + TJ) of padded planes [i0, i0 + CI) and streams through them:
// at step s it updates colour c1 of plane s on its rows plus one halo row on
// each side, keeping the new values in an LDS ring of three planes, then
// updates colour c2 of plane s - 1 from the ring.  Colour c1 is updated in
// place (no cell reads c1's old value); colour c2 is read from xBin and
// written to xBout, a second buffer, so a halo update of c1 always sees the
// old c2 whatever the neighbouring workgroups have written.  Outside rg,
// c1 keeps its values and c2 is copied to xBout, so xBout is a complete
// colour array.  Between the first colour's HBM reads (its b, coefficients
// and the old c2) and the second's, every coefficient and right-hand-side
// value is read once per launch instead of once per phase.
struct FusedTiles {
    int TJ, CI;        // rows and planes owned by a workgroup
    int nJt, nIt;      // tiles along J and I
    int per;           // tiles per XCD
    int bs;            // threads per workgroup (256, 512 or 1024)
    unsigned magic;    // ceil(2^32 / Hz): e / Hz = umulhi(e, magic)
};


template <typename T, int OP>
__device__ __forceinline__ T fused_update(const LevelDims& L,
                                          const OpView<T>& op, int colour,
                                          unsigned idx, int row, int p,
                                          int q, int I, int J, int K,
                                          const T nb[6], T bv)
{
    T a[7];
    coeffs<T, OP>(L, op, colour, int(idx), row, p, q, I, J, K, a);
    return gs_update<T>(a, nb, bv);
}

__device__ __forceinline__ bool in_region(const Region& rg, int I, int J,
                                          int K)
{
    return I >= rg.S[0] && I < rg.E[0] && J >= rg.S[1] && J < rg.E[1] &&
           K >= rg.S[2] && K < rg.E[2];
}

template <typename T, int OP, int BS>
__device__ __forceinline__ void rbgs2_body(const Bid bid, const LevelDims& L,
                                           const OpView<T>& op,
                                           T* __restrict__ xR,
                                           const T* __restrict__ xBin,
                                           T* __restrict__ xBout,
                                           const T* __restrict__ bR,
                                           const T* __restrict__ bB, int c1,
                                           const Region& rg,
                                           const FusedTiles& ft, T* ring)
{
    const unsigned lin = (bid.x & 7u)*unsigned(ft.per) + (bid.x >> 3);
    if (lin >= unsigned(ft.nJt*ft.nIt)) return;
    const int jt = int(lin % unsigned(ft.nJt));
    const int it = int(lin/unsigned(ft.nJt));
    const int P0 = L.P[0], P1 = L.P[1], P2 = L.P[2], Hz = L.Hz;
    const unsigned sI = unsigned(L.sI), uHz = unsigned(Hz);
    const int j0 = jt*ft.TJ, j1 = min(j0 + ft.TJ, P1);
    const int i0 = it*ft.CI, i1 = min(i0 + ft.CI, P0);
    // ring rows j0-1 .. j1 (row j0-1 at ring row 0), four plane slots
    const int rowsR = j1 - j0 + 2;
    const int slotSize = rowsR*Hz;
    const int c2 = c1 ^ 1;
    const int J0 = max(j0 - 1, 0), J1 = min(j1 + 1, P1);
    // step s: colour c1 of plane s into slot s % 4 and, in the same pass,
    // colour c2 of plane s - 2 from the slots of planes s-3, s-2, s-1
    for (int s = i0 - 1; s <= i1 + 1; s++)
    {
        const bool doR = s >= 0 && s < P0 && s <= i1;
        const int sB = s - 2;
        const bool doB = sB >= i0 && sB < i1;
        const int nR = doR ? (J1 - J0)*Hz : 0;
        const int nB = doB ? (j1 - j0)*Hz : 0;
        for (int e0 = int(threadIdx.x); e0 < nR + nB; e0 += BS)
        {
            if (e0 < nR)
            {
                const int e = e0;
                const int jr = int(__umulhi(unsigned(e), ft.magic));
                const int p = e - jr*Hz;
                const int J = J0 + jr;
                const int q = (c1 - s - J) & 1;
                const int K = 2*p + q;
                if (K >= P2) continue;
                const int row = (s*P1 + J)*Hz;
                const unsigned idx = unsigned(row + p);
                T v;
                if (in_region(rg, s, J, K))
                {
                    const T nb[6] = {ld(xBin, idx - sI), ld(xBin, idx + sI),
                                     ld(xBin, idx - uHz), ld(xBin, idx + uHz),
                                     ld(xBin, idx - 1 + q), ld(xBin, idx + q)};
                    const T bv = ld(bR, idx);
                    v = fused_update<T, OP>(L, op, c1, idx, row, p, q, s, J,
                                            K, nb, bv);
                    if (s >= i0 && s < i1 && J >= j0 && J < j1)
                        st(xR, idx, v);
                }
                else
                    v = ld(xR, idx);
                ring[(s & 3)*slotSize + (J - j0 + 1)*Hz + p] = v;
            }
            else
            {
                const int e = e0 - nR;
                const int jr = int(__umulhi(unsigned(e), ft.magic));
                const int p = e - jr*Hz;
                const int J = j0 + jr;
                const int q = (c2 - sB - J) & 1;
                const int K = 2*p + q;
                if (K >= P2) continue;
                const int row = (sB*P1 + J)*Hz;
                const unsigned idx = unsigned(row + p);
                if (in_region(rg, sB, J, K))
                {
                    const T* rm = ring + ((sB - 1) & 3)*slotSize;
                    const T* r0 = ring + (sB & 3)*slotSize;
                    const T* rp = ring + ((sB + 1) & 3)*slotSize;
                    const int lr = (jr + 1)*Hz + p;
                    const T nb[6] = {rm[lr], rp[lr], r0[lr - Hz], r0[lr + Hz],
                                     r0[lr - 1 + q], r0[lr + q]};
                    const T bv = ld(bB, idx);
                    st(xBout, idx, fused_update<T, OP>(L, op, c2, idx, row, p,
                                                       q, sB, J, K, nb, bv));
                }
                else
                    st(xBout, idx, ld(xBin, idx));
            }
        }
        __syncthreads();
    }
}

template <typename T, int OP, int BS>
__global__ __launch_bounds__(BS) void k_rbgs2(LevelDims L, OpView<T> op,
                                                    T* __restrict__ xR,
                                                    const T* __restrict__ xBin,
                                                    T* __restrict__ xBout,
                                                    const T* __restrict__ bR,
                                                    const T* __restrict__ bB,
                                                    int c1, Region rg,
                                                    FusedTiles ft)
{
    extern __shared__ unsigned char fused_lds[];
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    rbgs2_body<T, OP, BS>(bid, L, op, xR, xBin, xBout, bR, bB, c1, rg, ft,
                          reinterpret_cast<T*>(fused_lds));
}


// All red-black phases of the width-1 boundary slabs of one face pair
// (rbgsSmoother::SweepBoundary, rbgsSmoother.C:186-374) in one launch.  A
// slab is a plane of cells whose out-of-plane neighbours (the interior layer
// and the ghost layer filled just before) stay fixed while its phases run,
// so every phase is a 2-D red-black update of the plane.
//
// A workgroup owns a TT x TT tile (or the whole plane when it fits in
// 32 x 32 with its ghost ring), loads the tile grown by H cells (H = number
// of phases) into LDS and runs the phases there: phase p updates the tile
// grown by H-1-p cells, whose neighbours are exact after phase p-1, so the
// tile is exact after the last phase and the halo work is redundant but
// identical.  The in-plane values are read from a snapshot of the plane
// taken when the ghosts were filled (or written into it by an earlier slab
// launch), never from x, which the neighbouring workgroups overwrite while
// this one runs.  Everything else a cell needs (b, its two out-of-plane
// neighbours, its coefficients) is loaded once before the first phase.
// Opposite slabs of a pair are independent (they are at least two cells
// apart) and run as blockIdx.z = 0/1.  Results go to x and to the snapshots
// of the slab launches that follow (tptr/tdim/tplane), whose planes meet
// this one along a line.
template <typename T>
struct SlabArgs {
    int plane[2];      // padded index of each slab plane
    int nphase;        // H
    int colours[2][4];
    int tt;            // output tile edge; 0: one tile holds the whole plane
    const T* src[2];   // snapshot of each slab plane, P_a1 x P_a2
    int ntarget;
    T* tptr[6];
    int tdim[6];
    int tplane[6];
    int axis;          // D, the axis normal to the slabs
};

constexpr int SLAB_W = 32;   // LDS tile edge

template <typename T, int OP, int D>
__device__ __forceinline__ void rbgs_slabs_body(const Bid bid,
                                                const LevelDims& L,
                                                const OpView<T>& op,
                                                Pair<T> x,
                                                Pair<const T> b,
                                                const SlabArgs<T>& sa,
                                                T (*xs)[SLAB_W + 1])
{
    constexpr int A1 = D == 0 ? 1 : 0;
    constexpr int A2 = D == 2 ? 1 : 2;
    const int slab = bid.z;
    const int H = sa.nphase;
    const int n1 = L.N[A1], n2 = L.N[A2];
    const bool whole = sa.tt == 0;
    const int W1 = whole ? n1 + 2 : sa.tt + 2*H;
    const int W2 = whole ? n2 + 2 : sa.tt + 2*H;
    const int o1 = whole ? 0 : 1 + int(bid.y)*sa.tt - H;
    const int o2 = whole ? 0 : 1 + int(bid.x)*sa.tt - H;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const T* src = sa.src[slab];
    const int Ps = L.P[A2];

    // the four cells of this thread: u1 = ty + 16*(k>>1), u2 = tx + 16*(k&1)
    bool act[4];
    int colk[4];
    unsigned idxk[4];
    T a[4][7], bv[4], om[4], opl[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
        const int u1 = ty + 16*(k >> 1), u2 = tx + 16*(k & 1);
        const int c1 = o1 + u1, c2 = o2 + u2;
        const bool inTile = u1 < W1 && u2 < W2;
        T v = T(0);
        if (inTile && c1 >= 0 && c1 < L.P[A1] && c2 >= 0 && c2 < Ps)
            v = src[c1*Ps + c2];
        if (inTile) xs[u1][u2] = v;
        act[k] = inTile && c1 >= 1 && c1 <= n1 && c2 >= 1 && c2 <= n2;
        colk[k] = 0;
        idxk[k] = 0;
        bv[k] = om[k] = opl[k] = T(0);
#pragma unroll
        for (int m = 0; m < 7; m++) a[k][m] = T(0);
        if (act[k])
        {
            int g[3];
            g[D] = sa.plane[slab];
            g[A1] = c1;
            g[A2] = c2;
            const int I = g[0], J = g[1], K = g[2];
            const int colour = (I + J + K) & 1;
            const int q = K & 1, p = K >> 1;
            const int row = (I*L.P[1] + J)*L.Hz;
            const unsigned idx = unsigned(row + p);
            const T* xo = col(x, colour ^ 1);
            if (D == 0)
            {
                om[k] = ld(xo, idx - unsigned(L.sI));
                opl[k] = ld(xo, idx + unsigned(L.sI));
            }
            else if (D == 1)
            {
                om[k] = ld(xo, idx - unsigned(L.Hz));
                opl[k] = ld(xo, idx + unsigned(L.Hz));
            }
            else
            {
                om[k] = ld(xo, idx - 1 + q);
                opl[k] = ld(xo, idx + q);
            }
            bv[k] = ld(col(b, colour), idx);
            coeffs<T, OP>(L, op, colour, int(idx), row, p, q, I, J, K, a[k]);
            colk[k] = colour;
            idxk[k] = idx;
        }
    }
    __syncthreads();
    for (int ph = 0; ph < H; ph++)
    {
        const int colour = sa.colours[slab][ph];
        const int lo = whole ? 0 : ph + 1;
        const int hi1 = whole ? W1 : W1 - ph - 1;
        const int hi2 = whole ? W2 : W2 - ph - 1;
#pragma unroll
        for (int k = 0; k < 4; k++)
        {
            const int u1 = ty + 16*(k >> 1), u2 = tx + 16*(k & 1);
            if (!act[k] || colk[k] != colour || u1 < lo || u1 >= hi1 ||
                u2 < lo || u2 >= hi2)
                continue;
            // neighbour values in the reference order -i,+i,-j,+j,-k,+k
            T nb[6];
            if (D == 0)
            {
                nb[0] = om[k];
                nb[1] = opl[k];
                nb[2] = xs[u1 - 1][u2];
                nb[3] = xs[u1 + 1][u2];
                nb[4] = xs[u1][u2 - 1];
                nb[5] = xs[u1][u2 + 1];
            }
            else if (D == 1)
            {
                nb[0] = xs[u1 - 1][u2];
                nb[1] = xs[u1 + 1][u2];
                nb[2] = om[k];
                nb[3] = opl[k];
                nb[4] = xs[u1][u2 - 1];
                nb[5] = xs[u1][u2 + 1];
            }
            else
            {
                nb[0] = xs[u1 - 1][u2];
                nb[1] = xs[u1 + 1][u2];
                nb[2] = xs[u1][u2 - 1];
                nb[3] = xs[u1][u2 + 1];
                nb[4] = om[k];
                nb[5] = opl[k];
            }
            xs[u1][u2] = gs_update<T>(a[k], nb, bv[k]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
        const int u1 = ty + 16*(k >> 1), u2 = tx + 16*(k & 1);
        if (!act[k]) continue;
        if (!whole && (u1 < H || u1 >= H + sa.tt || u2 < H || u2 >= H + sa.tt))
            continue;
        const T v = xs[u1][u2];
        st(col(x, colk[k]), idxk[k], v);
        int g[3];
        g[D] = sa.plane[slab];
        g[A1] = o1 + u1;
        g[A2] = o2 + u2;
        for (int t = 0; t < sa.ntarget; t++)
        {
            const int d = sa.tdim[t];
            if (g[d] != sa.tplane[t]) continue;
            const int e1 = d == 0 ? g[1] : g[0];
            const int e2 = d == 2 ? g[1] : g[2];
            const int pe = d == 2 ? L.P[1] : L.P[2];
            sa.tptr[t][e1*pe + e2] = v;
        }
    }
}

template <typename T, int OP, int D>
__global__ __launch_bounds__(256) void k_rbgs_slabs(LevelDims L, OpView<T> op,
                                                    Pair<T> x, Pair<const T> b,
                                                    SlabArgs<T> sa)
{
    __shared__ T xs[SLAB_W][SLAB_W + 1];
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    rbgs_slabs_body<T, OP, D>(bid, L, op, x, b, sa, xs);
}

template <typename T, int OP, int D>
struct KSlabs
{
    static constexpr int bx = 256, by = 1;
    static constexpr bool lds = true;
    struct Args
    {
        LevelDims L;
        OpView<T> op;
        Pair<T> x;
        Pair<const T> b;
        SlabArgs<T> sa;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        __shared__ T xs[SLAB_W][SLAB_W + 1];
        rbgs_slabs_body<T, OP, D>(bid, a.L, a.op, a.x, a.b, a.sa, xs);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_rbgs_slabs<T, OP, D>), g, dim3(bx, by), shm, s,
                           a.L, a.op, a.x, a.b, a.sa);
    }
};


// Lexicographic Gauss-Seidel over region [S,E) (lexgsSmoother.C:14-124,
// SweepBoundary 126-254): a forward pass in (i,j,k) order, then a backward
// pass in the reverse order, each cell
//   x_c = (b_c - (A1 x_-i + A2 x_+i + A3 x_-j + A4 x_+j + A5 x_-k + A6 x_+k))
//         / A0.
// In the forward pass a cell reads its three lower neighbours after their
// update and its three upper neighbours before theirs (the reverse in the
// backward pass), and no two cells of one hyperplane (i-S0)+(j-S1)+(k-S2) =
// s are neighbours.  One launch per hyperplane, in pass order, therefore
// hands every cell the very values the sequential sweep hands it.  Block
// row y is i = S0 + ilo + y; its threads take the j of the hyperplane's
// (i, j) segment, k follows.
template <typename T, int OP>
__global__ __launch_bounds__(256) void k_lexgs_plane(LevelDims L,
                                                     OpView<T> op, Pair<T> x,
                                                     Pair<const T> b,
                                                     Region rg, int s, int ilo)
{
    const int M1 = rg.E[1] - rg.S[1], M2 = rg.E[2] - rg.S[2];
    const int ii = ilo + int(blockIdx.y);
    const int jlo = max(0, s - ii - (M2 - 1));
    const int jhi = min(M1 - 1, s - ii);
    const int jj = jlo + int(blockIdx.x*blockDim.x + threadIdx.x);
    if (jj > jhi) return;
    const int kk = s - ii - jj;
    if (kk < 0 || kk >= M2) return;
    const int I = rg.S[0] + ii, J = rg.S[1] + jj, K = rg.S[2] + kk;
    const int colour = (I + J + K) & 1;
    const int q = K & 1, p = K >> 1;
    const int row = (I*L.P[1] + J)*L.Hz;
    const unsigned idx = unsigned(row + p);
    const unsigned sI = L.sI, Hz = L.Hz;
    const T* xo = col(x, colour ^ 1);
    const T nb[6] = {ld(xo, idx - sI), ld(xo, idx + sI), ld(xo, idx - Hz),
                     ld(xo, idx + Hz), ld(xo, idx - 1 + q), ld(xo, idx + q)};
    const T bv = ld(col(b, colour), idx);
    T a[7];
    coeffs<T, OP>(L, op, colour, int(idx), row, p, q, I, J, K, a);
    st(col(x, colour), idx, gs_update<T>(a, nb, bv));
}

// Damped Jacobi (jacSmoother.C:18-127, omega 0.8): y from x, then copy back.
template <typename T, int OP>
__global__ __launch_bounds__(BX*BY) void k_jacobi(LevelDims L, OpView<T> op,
                                                  Pair<T> x, Pair<const T> b,
                                                  Pair<T> y, Region rg)
{
    const int I = rg.S[0] + blockIdx.z;
    const int J = rg.S[1] + blockIdx.y*BY + threadIdx.y;
    const int p = (rg.S[2] >> 1) + blockIdx.x*BX + threadIdx.x;
    if (J >= rg.E[1]) return;
    const int sI = L.sI;
    const int row = (I*L.P[1] + J)*L.Hz;
    const int idx = row + p;
    const T omega = T(0.8);
    const T om1 = T(1.0) - T(0.8);
#pragma unroll
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K < rg.S[2] || K >= rg.E[2]) continue;
        const int colour = (I + J + K) & 1;
        const T* xo = x.c[colour ^ 1];
        T a[7];
        coeffs<T, OP>(L, op, colour, idx, row, p, h, I, J, K, a);
        const T P =
            a[1]*xo[idx - sI]
          + a[2]*xo[idx + sI]
          + a[3]*xo[idx - L.Hz]
          + a[4]*xo[idx + L.Hz]
          + a[5]*xo[row + p - 1 + h]
          + a[6]*xo[row + p + h];
        y.c[colour][idx] = om1*x.c[colour][idx]
                         + omega*(b.c[colour][idx] - P)/a[0];
    }
}

template <typename T>
__global__ __launch_bounds__(BX*BY) void k_copy_region(LevelDims L,
                                                       Pair<const T> y,
                                                       Pair<T> x, Region rg)
{
    const int I = rg.S[0] + blockIdx.z;
    const int J = rg.S[1] + blockIdx.y*BY + threadIdx.y;
    const int p = (rg.S[2] >> 1) + blockIdx.x*BX + threadIdx.x;
    if (J >= rg.E[1]) return;
    const int idx = (I*L.P[1] + J)*L.Hz + p;
#pragma unroll
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K < rg.S[2] || K >= rg.E[2]) continue;
        const int colour = (I + J + K) & 1;
        x.c[colour][idx] = y.c[colour][idx];
    }
}

// ---------------------------------------------------------------------------
// Residual r = b - A x (linearSystemResidual.C:75-95; row product in
// C,L,R,B,T,A,F order, linearSystemFunctions.H:42-49) over the interior.
// Optionally writes per-tile partial sums of r^2 (deterministic order).
template <typename T, int OP, bool NORM, bool PACK>
__device__ __forceinline__ void residual_impl(const Bid bid,
                                              const LevelDims& L,
                                              const OpView<T>& op,
                                              Pair<const T> x,
                                              Pair<const T> b,
                                              Pair<T> r,
                                              const Tiles& tl,
                                              double* partials)
{
    __shared__ double sm[BX*BY/64];
    int bx, by, bz;
    if (!tile_of(tl, bid.x, bx, by, bz)) return;
    PairPos c;
    double acc = 0.0;
    if (pair_pos<PACK>(L, bx, by, bz, c, tl.ws))
    {
        const T bA = ld(col(b, c.cA), c.idx), bB = ld(col(b, c.cB), c.idx + 1);
        T Ax[2];
        pair_Ax<T, OP>(L, op, x, c, Ax);
        const T rA = bA - Ax[0], rB = bB - Ax[1];
        st(col(r, c.cA), c.idx, rA);
        if (NORM) acc += (double)rA*(double)rA;
        if (c.vB)
        {
            st(col(r, c.cB), c.idx + 1, rB);
            if (NORM) acc += (double)rB*(double)rB;
        }
    }
    if (NORM)
    {
        const double s = block_sum(acc, sm);
        if (threadIdx.x == 0 && threadIdx.y == 0)
            partials[tile_index(tl, bx, by, bz)] = s;
    }
}

// Rows of short levels are packed (tl.ws < 6) when no norm is taken; the
// norm's tiles keep one row segment per wavefront so that its partial sums
// keep their order.
template <typename T, int OP, bool NORM>
__device__ __forceinline__ void residual_body(const Bid bid,
                                              const LevelDims& L,
                                              const OpView<T>& op,
                                              Pair<const T> x,
                                              Pair<const T> b,
                                              Pair<T> r,
                                              const Tiles& tl,
                                              double* partials)
{
    if (!NORM && tl.ws != 6)
        residual_impl<T, OP, NORM, true>(bid, L, op, x, b, r, tl, partials);
    else
        residual_impl<T, OP, NORM, false>(bid, L, op, x, b, r, tl, partials);
}

template <typename T, int OP, bool NORM>
__global__ __launch_bounds__(BX*BY) void k_residual(LevelDims L, OpView<T> op,
                                                    Pair<const T> x,
                                                    Pair<const T> b, Pair<T> r,
                                                    Tiles tl, double* partials)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    residual_body<T, OP, NORM>(bid, L, op, x, b, r, tl, partials);
}

template <typename T, int OP, bool NORM>
struct KResidual
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = true;
    struct Args
    {
        LevelDims L;
        OpView<T> op;
        Pair<const T> x;
        Pair<const T> b;
        Pair<T> r;
        Tiles tl;
        double* partials;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        residual_body<T, OP, NORM>(bid, a.L, a.op, a.x, a.b, a.r, a.tl,
                                   a.partials);
    }
    // the norm's partial sums keep one tile shape (their summation order)
    static bool retile(Args& a, int ws)
    {
        if (NORM) return false;
        a.tl = packed_tiles((a.L.N[2] + 1)/2, a.L.N[1], a.L.N[0], ws);
        return true;
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_residual<T, OP, NORM>), g, dim3(bx, by), shm, s,
                           a.L, a.op, a.x, a.b, a.r, a.tl, a.partials);
    }
};


// 1/x: correctly rounded in FP64 (the reference's division); the hardware
// reciprocal in the FP32 cycles, which carry no bitwise contract.
__device__ __forceinline__ double recip(double x) { return 1.0/x; }
__device__ __forceinline__ float recip(float x)
{
    return __builtin_amdgcn_rcpf(x);
}

// c + a*b: product and sum each rounded in FP64 (the reference's order), one
// fused multiply-add in FP32.
__device__ __forceinline__ double madd(double a, double b, double c)
{
    return c + a*b;
}
__device__ __forceinline__ float madd(float a, float b, float c)
{
    return __builtin_fmaf(a, b, c);
}

// ---------------------------------------------------------------------------
// Scaled linear restriction (linearRestrictionScheme.C:103-160, weights
// :15-80): b_c = cvc * sum_q (w_q * r_q) * (1/cvf_q), q over the 8 children
// (z outer, x inner), w_q = ((fx*fy)*fz) zeroed when <= 1e-12.
//
// One thread per pair of coarse cells K = 2p+1, 2p+2 (cell-pair tiles of
// the coarse level); the 16 fine values are loaded before any arithmetic.
template <typename T, bool PACK>
__device__ __forceinline__ void restrict_impl(const Bid bid,
                                              const LevelDims& Lf,
                                              const LevelDims& Lc,
                                              const TransferView<T>& tv,
                                              Pair<const T> rf,
                                              Pair<T> bc,
                                              const Tiles& tl)
{
    int bx, by, bz;
    if (!tile_of(tl, bid.x, bx, by, bz)) return;
    PairPos c;
    if (!pair_pos<PACK>(Lc, bx, by, bz, c, tl.ws)) return;
    const int I = c.I, J = c.J;
    const int R0 = Lc.R[0], R1 = Lc.R[1], R2 = Lc.R[2];
    const int ox = R0 > 1, oy = R1 > 1, oz = R2 > 1;
    int fi[2], fj[2];
#pragma unroll
    for (int k = 0; k < 2; k++)
    {
        fi[k] = (I - 1)*R0 + k*ox + 1;
        fj[k] = (J - 1)*R1 + k*oy + 1;
    }
    // cell h = 0: K = 2p+1, h = 1: K = 2p+2 (a missing B reuses A's K)
    const int Kh[2] = {2*c.p + 1, c.vB ? 2*c.p + 2 : 2*c.p + 1};
    // fine values [h][cz][cy][cx]
    T rv[2][2][2][2];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
    for (int cz = 0; cz < 2; cz++)
#pragma unroll
    for (int cy = 0; cy < 2; cy++)
#pragma unroll
    for (int cx = 0; cx < 2; cx++)
    {
        const int fk = (Kh[h] - 1)*R2 + cz*oz + 1;
        rv[h][cz][cy][cx] = ld(col(rf, (fi[cx] + fj[cy] + fk) & 1),
                               unsigned(cidx(Lf, fi[cx], fj[cy], fk)));
    }
    // every weight factor loaded up front as well: no load under a branch,
    // so a term with w = 0 is dropped by a select (adding it would leave
    // the sum as it is for finite values; the select also for the others)
    T vf0[2], vf1[2], vf2[2][2], uz[2], cw2[2];
#pragma unroll
    for (int k = 0; k < 2; k++)
    {
        vf0[k] = tv.fw[0][fi[k]];
        vf1[k] = tv.fw[1][fj[k]];
        uz[k] = tv.ru[2][Kh[k]];
        cw2[k] = tv.cw[2][Kh[k]];
#pragma unroll
        for (int cz = 0; cz < 2; cz++)
            vf2[k][cz] = tv.fw[2][(Kh[k] - 1)*R2 + cz*oz + 1];
    }
    const T ux = tv.ru[0][I], uy = tv.ru[1][J];
    const T cw0 = tv.cw[0][I], cw1 = tv.cw[1][J];
    const T fx[2] = {T(1.0) - ux, ux};
    const T fy[2] = {T(1.0) - uy, uy};
    T out[2];
#pragma unroll
    for (int h = 0; h < 2; h++)
    {
        const T fz[2] = {T(1.0) - uz[h], uz[h]};
        T sum = T(0);
#pragma unroll
        for (int cz = 0; cz < 2; cz++)
#pragma unroll
        for (int cy = 0; cy < 2; cy++)
#pragma unroll
        for (int cx = 0; cx < 2; cx++)
        {
            T w = fx[cx]*fy[cy];
            w = w*fz[cz];
            if (w <= T(1e-12)) w = T(0);
            T cvf = vf0[cx]*vf1[cy];
            cvf = cvf*vf2[h][cz];
            cvf = cvf > T(1e-16) ? cvf : T(1e-16);
            const T icvf = recip(cvf);
            const T t = w*rv[h][cz][cy][cx];
            const T next = madd(t, icvf, sum);
            sum = w != T(0) ? next : sum;
        }
        T cvc = cw0*cw1;
        cvc = cvc*cw2[h];
        cvc = cvc > T(1e-16) ? cvc : T(1e-16);
        out[h] = sum*cvc;
    }
    st(col(bc, c.cA), c.idx, out[0]);
    if (c.vB) st(col(bc, c.cB), c.idx + 1, out[1]);
}

template <typename T>
__device__ __forceinline__ void restrict_body(const Bid bid,
                                              const LevelDims& Lf,
                                              const LevelDims& Lc,
                                              const TransferView<T>& tv,
                                              Pair<const T> rf,
                                              Pair<T> bc,
                                              const Tiles& tl)
{
    if (tl.ws != 6)
        restrict_impl<T, true>(bid, Lf, Lc, tv, rf, bc, tl);
    else
        restrict_impl<T, false>(bid, Lf, Lc, tv, rf, bc, tl);
}

template <typename T>
__global__ __launch_bounds__(BX*BY) void k_restrict(LevelDims Lf, LevelDims Lc,
                                                    TransferView<T> tv,
                                                    Pair<const T> rf,
                                                    Pair<T> bc, Tiles tl)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    restrict_body<T>(bid, Lf, Lc, tv, rf, bc, tl);
}

template <typename T>
struct KRestrict
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims Lf;
        LevelDims Lc;
        TransferView<T> tv;
        Pair<const T> rf;
        Pair<T> bc;
        Tiles tl;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        restrict_body<T>(bid, a.Lf, a.Lc, a.tv, a.rf, a.bc, a.tl);
    }
    static bool retile(Args& a, int ws)
    {
        a.tl = packed_tiles((a.Lc.N[2] + 1)/2, a.Lc.N[1], a.Lc.N[0], ws);
        return true;
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_restrict<T>), g, dim3(bx, by), shm, s, a.Lf,
                           a.Lc, a.tv, a.rf, a.bc, a.tl);
    }
};


// Linear prolongation with plusEqOp (linearProlongationScheme.C:112-147,
// weights :12-110): x_f += sum_v w_v x_c(idx_v), v = (cx,cy,cz) with x
// fastest, idx = i/R + abc*off, off = i odd ? +1 : -1, w_v = (gx*gy)*gz with
// g = 1-u (abc 0) or u (abc 1); u = 1/2 on all axes when any u lies outside
// [0,1) (:89-93).
//
// One thread per pair of fine cells K = 2p+1 (k = 2p even: neighbour
// below) and K = 2p+2 (k = 2p+1 odd: neighbour above).  Their coarse z
// indices lie in zbase..zbase+nz-1 (R = 2: parents p+1, neighbours p and
// p+2; R = 1: cells 2p+1, 2p+2 with neighbours 2p and 2p+3), loaded once
// for the 4 coarse columns.
template <typename T, bool PACK>
__device__ __forceinline__ void prolong_impl(const Bid bid,
                                             const LevelDims& Lf,
                                             const LevelDims& Lc,
                                             const TransferView<T>& tv,
                                             Pair<const T> xc,
                                             Pair<T> xf,
                                             const Tiles& tl)
{
    int bx, by, bz;
    if (!tile_of(tl, bid.x, bx, by, bz)) return;
    PairPos c;
    if (!pair_pos<PACK>(Lf, bx, by, bz, c, tl.ws)) return;
    const int I = c.I, J = c.J, p = c.p;
    const int i = I - 1, j = J - 1;
    const int offx = (i & 1) ? 1 : -1, offy = (j & 1) ? 1 : -1;
    const int ci[2] = {i/Lc.R[0] + 1, i/Lc.R[0] + 1 + offx};
    const int cj[2] = {j/Lc.R[1] + 1, j/Lc.R[1] + 1 + offy};
    const int R2 = Lc.R[2];
    const int zbase = R2 == 1 ? 2*p : p;
    const int nz = R2 == 1 ? 4 : 3;
    // coarse values [z - zbase][cy][cx]
    T cv[4][2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
    for (int bb = 0; bb < 2; bb++)
    {
        const int base = (ci[a]*Lc.P[1] + cj[bb])*Lc.Hz;
        const int par = ci[a] + cj[bb];
#pragma unroll
        for (int z = 0; z < 4; z++)
        {
            // for nz = 3 the fourth value repeats the third and is unused
            const int zz = zbase + (z < nz ? z : nz - 1);
            cv[z][bb][a] = ld(col(xc, (par + zz) & 1),
                               unsigned(base + (zz >> 1)));
        }
    }
    const T old[2] = {ld(col(xf, c.cA), c.idx), ld(col(xf, c.cB), c.idx + 1)};
    const T ux = tv.pu[0][I], uy = tv.pu[1][J];
    T add[2] = {T(0), T(0)};
#pragma unroll
    for (int h = 0; h < 2; h++)
    {
        if (h == 1 && !c.vB) continue;
        const int K = 2*p + 1 + h;
        // parent and neighbour as indices into cv: cell A (k = 2p even)
        // has parent 1, neighbour 0; cell B (k = 2p+1 odd) parent 1,
        // neighbour 2 for R = 2, parent 2, neighbour 3 for R = 1
        const int zpar = h == 0 ? 1 : (R2 == 1 ? 2 : 1);
        const int znb = h == 0 ? 0 : zpar + 1;
        T u0 = ux, u1 = uy, u2 = tv.pu[2][K];
        const bool inside = !(u0 < T(0) || u0 >= T(1) || u1 < T(0) ||
                              u1 >= T(1) || u2 < T(0) || u2 >= T(1));
        if (!inside) { u0 = T(0.5); u1 = T(0.5); u2 = T(0.5); }
        const T gx[2] = {T(1.0) - u0, u0};
        const T gy[2] = {T(1.0) - u1, u1};
        const T gz[2] = {T(1.0) - u2, u2};
        const int zsel[2] = {zpar, znb};
        // the reference skips zero weights (linearProlongationScheme.C:
        // 139-143); adding w*c = +-0 for w = 0 leaves the sum unchanged for
        // finite c, so every term is added
        T value = T(0);
#pragma unroll
        for (int cz = 0; cz < 2; cz++)
#pragma unroll
        for (int cy = 0; cy < 2; cy++)
#pragma unroll
        for (int cx = 0; cx < 2; cx++)
        {
            T w = gx[cx]*gy[cy];
            w = w*gz[cz];
            const T t = w*cv[zsel[cz]][cy][cx];
            value = value + t;
        }
        add[h] = value;
    }
    st(col(xf, c.cA), c.idx, old[0] + add[0]);
    if (c.vB) st(col(xf, c.cB), c.idx + 1, old[1] + add[1]);
}

template <typename T>
__device__ __forceinline__ void prolong_body(const Bid bid,
                                             const LevelDims& Lf,
                                             const LevelDims& Lc,
                                             const TransferView<T>& tv,
                                             Pair<const T> xc,
                                             Pair<T> xf,
                                             const Tiles& tl)
{
    if (tl.ws != 6)
        prolong_impl<T, true>(bid, Lf, Lc, tv, xc, xf, tl);
    else
        prolong_impl<T, false>(bid, Lf, Lc, tv, xc, xf, tl);
}

template <typename T>
__global__ __launch_bounds__(BX*BY) void k_prolong(LevelDims Lf, LevelDims Lc,
                                                   TransferView<T> tv,
                                                   Pair<const T> xc,
                                                   Pair<T> xf, Tiles tl)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    prolong_body<T>(bid, Lf, Lc, tv, xc, xf, tl);
}

template <typename T>
struct KProlong
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims Lf;
        LevelDims Lc;
        TransferView<T> tv;
        Pair<const T> xc;
        Pair<T> xf;
        Tiles tl;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        prolong_body<T>(bid, a.Lf, a.Lc, a.tv, a.xc, a.xf, a.tl);
    }
    static bool retile(Args& a, int ws)
    {
        a.tl = packed_tiles((a.Lf.N[2] + 1)/2, a.Lf.N[1], a.Lf.N[0], ws);
        return true;
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_prolong<T>), g, dim3(bx, by), shm, s, a.Lf, a.Lc,
                           a.tv, a.xc, a.xf, a.tl);
    }
};

// The same prolongation organised by coarse cell, for levels refined by 2
// along every axis (N_f = 2 N_c): thread (ci, cj, ck) adds into its eight
// children I = 2ci-1+a, J = 2cj-1+b, K = 2ck-1+c.  A child with offset a
// (fine interior index i = 2(ci-1)+a) has parent ci and neighbour
// ci-1 (a = 0, i even) or ci+1 (a = 1, i odd), so the 27 coarse values of
// the 3x3x3 block around (ci, cj, ck) serve all eight children: 27 coarse
// loads per 8 fine cells instead of 8 per cell.  Lanes run along ck, so the
// children K = 2ck-1 (colour (I+J+1)&1, position ck-1) and K = 2ck (colour
// (I+J)&1, position ck) of one fine row are contiguous in one colour array
// across the wavefront.  Per child the arithmetic is prolong_impl's.
template <typename T, bool PACK>
__device__ __forceinline__ void prolong8_impl(const Bid bid,
                                              const LevelDims& Lf,
                                              const LevelDims& Lc,
                                              const TransferView<T>& tv,
                                              Pair<const T> xc,
                                              Pair<T> xf,
                                              const Tiles& tl, int skip)
{
    int bx, by, bz;
    if (!tile_of(tl, bid.x, bx, by, bz)) return;
    const int ci = 1 + bz;
    int cj, ck;
    if (PACK)
    {
        const int lane = int(threadIdx.x), ws = tl.ws;
        cj = 1 + (by*BY + int(threadIdx.y))*(64 >> ws) + (lane >> ws);
        ck = 1 + (bx << ws) + (lane & ((1 << ws) - 1));
    }
    else
    {
        cj = __builtin_amdgcn_readfirstlane(1 + by*BY + int(threadIdx.y));
        ck = 1 + bx*BX + int(threadIdx.x);
    }
    if (cj > Lc.N[1] || ck > Lc.N[2]) return;
    // the children first: K = 2ck-1 (c = 0) and K = 2ck (c = 1) of the
    // fine rows (I, J) = (2ci-1+a, 2cj-1+b); children of colour `skip` are
    // left as they are (the caller's next phase overwrites every cell of
    // that colour).  Their loads are issued before the coarse block's, so
    // the memory latency of both overlaps.
    T old[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
    {
        const int I = 2*ci - 1 + a, J = 2*cj - 1 + b;
        const int cB = (I + J) & 1;
        const unsigned row = unsigned((I*Lf.P[1] + J)*Lf.Hz);
        old[a][b][0] = (cB ^ 1) != skip
                           ? ld(col(xf, cB ^ 1), row + unsigned(ck - 1)) : T(0);
        old[a][b][1] = cB != skip ? ld(col(xf, cB), row + unsigned(ck)) : T(0);
    }
    // coarse block [x][y][z] = (ci-1+x, cj-1+y, ck-1+z)
    T cv[3][3][3];
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
    for (int y = 0; y < 3; y++)
    {
        const int X = ci - 1 + x, Y = cj - 1 + y;
        const int base = (X*Lc.P[1] + Y)*Lc.Hz;
#pragma unroll
        for (int z = 0; z < 3; z++)
        {
            const int Z = ck - 1 + z;
            cv[x][y][z] = ld(col(xc, (X + Y + Z) & 1), unsigned(base + (Z >> 1)));
        }
    }
    // prolongation fractions of the children along each axis
    T uX[2], uY[2], uZ[2];
#pragma unroll
    for (int a = 0; a < 2; a++)
    {
        uX[a] = tv.pu[0][2*ci - 1 + a];
        uY[a] = tv.pu[1][2*cj - 1 + a];
        uZ[a] = tv.pu[2][2*ck - 1 + a];
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
    {
        const int I = 2*ci - 1 + a, J = 2*cj - 1 + b;
        const int cB = (I + J) & 1;
        const unsigned row = unsigned((I*Lf.P[1] + J)*Lf.Hz);
        T* xa = col(xf, cB ^ 1);
        T* xb = col(xf, cB);
        const unsigned ia = row + unsigned(ck - 1), ib = row + unsigned(ck);
        const bool doA = (cB ^ 1) != skip, doB = cB != skip;
        // parent (abc 0) and neighbour (abc 1) positions in the block:
        // parent 1, neighbour 0 (offset 0) or 2 (offset 1)
        const int sx[2] = {1, a ? 2 : 0};
        const int sy[2] = {1, b ? 2 : 0};
        T add[2];
#pragma unroll
        for (int c = 0; c < 2; c++)
        {
            const int sz[2] = {1, c ? 2 : 0};
            T u0 = uX[a], u1 = uY[b], u2 = uZ[c];
            const bool inside = !(u0 < T(0) || u0 >= T(1) || u1 < T(0) ||
                                  u1 >= T(1) || u2 < T(0) || u2 >= T(1));
            if (!inside) { u0 = T(0.5); u1 = T(0.5); u2 = T(0.5); }
            T value = T(0);
            if constexpr (std::is_same<T, float>::value)
            {
                // FP32 cycles: the same trilinear value as nested linear
                // interpolations in x, then y, then z (7 per child)
                float ry[2];
#pragma unroll
                for (int cz = 0; cz < 2; cz++)
                {
                    float rx[2];
#pragma unroll
                    for (int cy = 0; cy < 2; cy++)
                    {
                        const float v0 = cv[sx[0]][sy[cy]][sz[cz]];
                        const float v1 = cv[sx[1]][sy[cy]][sz[cz]];
                        rx[cy] = __builtin_fmaf(u0, v1 - v0, v0);
                    }
                    ry[cz] = __builtin_fmaf(u1, rx[1] - rx[0], rx[0]);
                }
                value = __builtin_fmaf(u2, ry[1] - ry[0], ry[0]);
            }
            else
            {
                const T gx[2] = {T(1.0) - u0, u0};
                const T gy[2] = {T(1.0) - u1, u1};
                const T gz[2] = {T(1.0) - u2, u2};
#pragma unroll
                for (int cz = 0; cz < 2; cz++)
#pragma unroll
                for (int cy = 0; cy < 2; cy++)
#pragma unroll
                for (int cx = 0; cx < 2; cx++)
                {
                    T w = gx[cx]*gy[cy];
                    w = w*gz[cz];
                    const T t = w*cv[sx[cx]][sy[cy]][sz[cz]];
                    value = value + t;
                }
            }
            add[c] = value;
        }
        if (doA) st(xa, ia, old[a][b][0] + add[0]);
        if (doB) st(xb, ib, old[a][b][1] + add[1]);
    }
}

template <typename T>
__global__ __launch_bounds__(BX*BY) void k_prolong8(LevelDims Lf, LevelDims Lc,
                                                    TransferView<T> tv,
                                                    Pair<const T> xc,
                                                    Pair<T> xf, Tiles tl,
                                                    int skip)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    if (tl.ws != 6)
        prolong8_impl<T, true>(bid, Lf, Lc, tv, xc, xf, tl, skip);
    else
        prolong8_impl<T, false>(bid, Lf, Lc, tv, xc, xf, tl, skip);
}

template <typename T>
struct KProlong8
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims Lf;
        LevelDims Lc;
        TransferView<T> tv;
        Pair<const T> xc;
        Pair<T> xf;
        Tiles tl;
        int skip;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        if (a.tl.ws != 6)
            prolong8_impl<T, true>(bid, a.Lf, a.Lc, a.tv, a.xc, a.xf, a.tl,
                                   a.skip);
        else
            prolong8_impl<T, false>(bid, a.Lf, a.Lc, a.tv, a.xc, a.xf, a.tl,
                                    a.skip);
    }
    static bool retile(Args& a, int ws)
    {
        a.tl = packed_tiles(a.Lc.N[2], a.Lc.N[1], a.Lc.N[0], ws);
        return true;
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_prolong8<T>), g, dim3(bx, by), shm, s, a.Lf,
                           a.Lc, a.tv, a.xc, a.xf, a.tl, a.skip);
    }
};


// ---------------------------------------------------------------------------
// Ghost values (meshLevel.C:231-326).  mode 0: periodic face ghosts only
// (evaluate<nonEliminatedBcs,1>); mode 2: Neumann/Dirichlet face ghosts only
// (correct<eliminatedBcs>, MG.C:475); mode 1: every ghost cell -- face ghosts by
// their boundary condition (periodic copy, NeumannBoundaryCondition.C:76-84,
// DirichletBoundaryCondition.C:64-70), edge/corner ghosts by periodic copy
// when every face they border is periodic, else a copy of the adjacent
// interior cell.  Every rule reads interior cells only, so the order in
// which ghosts are written does not matter.
template <typename T>
__device__ __forceinline__ void ghost_cell(const LevelDims& L,
                                           const OpView<T>& op, Pair<T> x,
                                           int mode, int level0,
                                           const BcVals& bv, int I, int J,
                                           int K)
{
    const int g[3] = {I, J, K};
    int bo[3];
    int nout = 0;
    for (int d = 0; d < 3; d++)
    {
        bo[d] = g[d] == 0 ? -1 : (g[d] == L.P[d] - 1 ? 1 : 0);
        nout += bo[d] != 0;
    }
    if (nout == 0) return;
    bool allPer = true;
    for (int d = 0; d < 3; d++)
        if (bo[d] != 0 && L.bc[2*d + (bo[d] > 0)] != PER) allPer = false;
    if (mode == 0 && !(nout == 1 && allPer)) return;
    if (mode == 2 && !(nout == 1 && !allPer)) return;
    int s[3];
    T val;
    if (allPer)
    {
        for (int d = 0; d < 3; d++) s[d] = g[d] - bo[d]*L.N[d];
        val = col(x, (s[0] + s[1] + s[2]) & 1)[cidx(L, s[0], s[1], s[2])];
    }
    else
    {
        for (int d = 0; d < 3; d++) s[d] = g[d] - bo[d];
        const T xi = col(x, (s[0] + s[1] + s[2]) & 1)[cidx(L, s[0], s[1], s[2])];
        if (nout > 1)
            val = xi;
        else
        {
            int d = bo[0] != 0 ? 0 : (bo[1] != 0 ? 1 : 2);
            const int f = 2*d + (bo[d] > 0);
            const T H = T(level0 ? 1.0 : 0.0);
            const T bvf = T(level0 ? bc_value(L, bv, f, s) : 0.0);
            if (L.bc[f] == NEU)
            {
                const T delta = op.dl[d][bo[d] > 0 ? L.N[d] + 1 : 1];
                val = xi + H*bvf/delta;
            }
            else
                val = H*T(2.0)*bvf - xi;
        }
    }
    col(x, (I + J + K) & 1)[cidx(L, I, J, K)] = val;
}

// The ghost shell as the coarse solve walks it: flat 32-bit index, K
// fastest.  A: planes I = 0 and I = P0-1; B: rows J = 0 and J = P1-1 of
// planes I = 1..P0-2; C: cells K = 0 and K = P2-1 of the remaining rows.
struct Shell {
    int nA, nB, nC;
};

__device__ __forceinline__ bool shell_cell(const LevelDims& L,
                                           const Shell& sh, int t, int& I,
                                           int& J, int& K)
{
    const int P0 = L.P[0], P1 = L.P[1], P2 = L.P[2];
    if (t < sh.nA)
    {
        const int half = P1*P2;
        const int side = t >= half;
        const int u = t - side*half;
        I = side ? P0 - 1 : 0;
        J = u/P2;
        K = u - J*P2;
        return true;
    }
    t -= sh.nA;
    if (t < sh.nB)
    {
        const int half = (P0 - 2)*P2;
        const int side = t >= half;
        const int u = t - side*half;
        J = side ? P1 - 1 : 0;
        const int i = u/P2;
        I = 1 + i;
        K = u - i*P2;
        return true;
    }
    t -= sh.nB;
    if (t < sh.nC)
    {
        const int half = (P0 - 2)*(P1 - 2);
        const int side = t >= half;
        const int u = t - side*half;
        K = side ? P2 - 1 : 0;
        const int i = u/(P1 - 2);
        I = 1 + i;
        J = 1 + (u - i*(P1 - 2));
        return true;
    }
    return false;
}

__host__ __device__ inline Shell shell_all(const LevelDims& L)
{
    const int P0 = L.P[0], P1 = L.P[1], P2 = L.P[2];
    Shell sh;
    sh.nA = 2*P1*P2;
    sh.nB = 2*(P0 - 2)*P2;
    sh.nC = 2*(P0 - 2)*(P1 - 2);
    return sh;
}

// Face ghost of face 2D+side at in-face interior cell (c1, c2) of the two
// other axes, by the rules of ghost_cell with every index static.
template <typename T, int D>
__device__ __forceinline__ void face_ghost(const LevelDims& L,
                                           const OpView<T>& op, Pair<T> x,
                                           int mode, int level0,
                                           const BcVals& bv, int side, int c1,
                                           int c2)
{
    constexpr int A1 = D == 0 ? 1 : 0;
    constexpr int A2 = D == 2 ? 1 : 2;
    const int bcf = side ? L.bc[2*D + 1] : L.bc[2*D];
    const bool per = bcf == PER;
    if (mode == 0 && !per) return;
    if (mode == 2 && per) return;
    const int N = L.N[D];
    int g[3], s[3];
    g[D] = side ? N + 1 : 0;
    g[A1] = c1;
    g[A2] = c2;
    s[A1] = c1;
    s[A2] = c2;
    s[D] = per ? (side ? 1 : N) : (side ? N : 1);
    const T xs = ld(col(x, (s[0] + s[1] + s[2]) & 1),
                    unsigned(cidx(L, s[0], s[1], s[2])));
    T val;
    if (per)
        val = xs;
    else
    {
        const T H = T(level0 ? 1.0 : 0.0);
        const T bvf = T(level0 ? bc_value(L, bv, 2*D + side, s) : 0.0);
        if (bcf == NEU)
        {
            const T delta = ld(op.dl[D], unsigned(side ? N + 1 : 1));
            val = xs + H*bvf/delta;
        }
        else
            val = H*T(2.0)*bvf - xs;
    }
    st(col(x, (g[0] + g[1] + g[2]) & 1), unsigned(cidx(L, g[0], g[1], g[2])),
       val);
}

// Edge or corner ghost (two or three indices outside): periodic image when
// every face it borders is periodic, else a copy of the adjacent interior
// cell.
template <typename T>
__device__ __forceinline__ void edge_ghost(const LevelDims& L, Pair<T> x,
                                           int I, int J, int K)
{
    const int g[3] = {I, J, K};
    int bo[3];
    bool allPer = true;
#pragma unroll
    for (int d = 0; d < 3; d++)
    {
        bo[d] = g[d] == 0 ? -1 : (g[d] == L.P[d] - 1 ? 1 : 0);
        const int bcf = bo[d] < 0 ? L.bc[2*d] : L.bc[2*d + 1];
        if (bo[d] != 0 && bcf != PER) allPer = false;
    }
    int s[3];
#pragma unroll
    for (int d = 0; d < 3; d++)
        s[d] = allPer ? g[d] - bo[d]*L.N[d] : g[d] - bo[d];
    const T v = ld(col(x, (s[0] + s[1] + s[2]) & 1),
                   unsigned(cidx(L, s[0], s[1], s[2])));
    st(col(x, (I + J + K) & 1), unsigned(cidx(L, I, J, K)), v);
}

// Ghost fill in up to four groups of threads, each padded to whole blocks
// so that a block serves one group: group D < 3 the two faces normal to
// axis D (face ghosts only), group 3 the edges and corners (mode 1).
struct BcGroups {
    int start[8];     // first thread of each group; start[7] = total
    int count[7];     // cells in each group (0: not filled)
};

// Snapshot cell (c1, c2) of the slab plane of axis D in slot `slot` (plane 1
// or N_D) as x holds it once the periodic face ghosts are filled: an
// in-plane face ghost of a periodic face takes its periodic image, every
// other cell its current value.  Only cells this kernel never writes are
// read, so the snapshot does not depend on the order of the threads.
template <typename T, int D>
__device__ __forceinline__ void snap_cell(const LevelDims& L, Pair<T> x,
                                          int slot, int c1, int c2, T* dst)
{
    constexpr int A1 = D == 0 ? 1 : 0;
    constexpr int A2 = D == 2 ? 1 : 2;
    int g[3];
    g[D] = slot ? L.N[D] : 1;
    g[A1] = c1;
    g[A2] = c2;
    const bool out1 = c1 == 0 || c1 == L.P[A1] - 1;
    const bool out2 = c2 == 0 || c2 == L.P[A2] - 1;
    if (out1 && !out2)
    {
        const int bcf = c1 == 0 ? L.bc[2*A1] : L.bc[2*A1 + 1];
        if (bcf == PER) g[A1] = c1 == 0 ? L.N[A1] : 1;
    }
    else if (out2 && !out1)
    {
        const int bcf = c2 == 0 ? L.bc[2*A2] : L.bc[2*A2 + 1];
        if (bcf == PER) g[A2] = c2 == 0 ? L.N[A2] : 1;
    }
    dst[c1*L.P[A2] + c2] = ld(col(x, (g[0] + g[1] + g[2]) & 1),
                              unsigned(cidx(L, g[0], g[1], g[2])));
}

template <typename T>
struct SnapPtrs {
    T* p[3][2];
};

template <typename T>
__device__ __forceinline__ void bc_fill_body(const Bid bid, const LevelDims& L,
                                             const OpView<T>& op,
                                             Pair<T> x, int mode,
                                             int level0, const BcVals& bv,
                                             const BcGroups& gr,
                                             const SnapPtrs<T>& sn,
                                             int tid = -1)
{
    // tid: the thread's index among the 256 of the workgroup (threadIdx.x
    // by default; the merged phase-and-fill launch passes its own)
    const int b0 = bid.x*256;
    int g = 0;
    while (g < 6 && b0 >= gr.start[g + 1]) g++;
    const int t = b0 + (tid >= 0 ? tid : int(threadIdx.x)) - gr.start[g];
    if (t >= gr.count[g]) return;
    if (g >= 4)
    {
        // slab-plane snapshots of axis g-4, both slots
        const int d = g - 4;
        const int p1 = d == 0 ? L.P[1] : L.P[0];
        const int p2 = d == 2 ? L.P[1] : L.P[2];
        const int slot = t >= p1*p2;
        const int u = t - slot*p1*p2;
        const int c1 = u/p2;
        const int c2 = u - c1*p2;
        if (d == 0) snap_cell<T, 0>(L, x, slot, c1, c2, sn.p[0][slot]);
        else if (d == 1) snap_cell<T, 1>(L, x, slot, c1, c2, sn.p[1][slot]);
        else snap_cell<T, 2>(L, x, slot, c1, c2, sn.p[2][slot]);
        return;
    }
    if (g < 3)
    {
        const int n1 = g == 0 ? L.N[1] : L.N[0];
        const int n2 = g == 2 ? L.N[1] : L.N[2];
        const int half = n1*n2;
        const int side = t >= half;
        const int u = t - side*half;
        const int c1 = u/n2;
        const int c2 = u - c1*n2;
        if (g == 0)
            face_ghost<T, 0>(L, op, x, mode, level0, bv, side, 1 + c1, 1 + c2);
        else if (g == 1)
            face_ghost<T, 1>(L, op, x, mode, level0, bv, side, 1 + c1, 1 + c2);
        else
            face_ghost<T, 2>(L, op, x, mode, level0, bv, side, 1 + c1, 1 + c2);
        return;
    }
    // edges along axis 0, 1, 2 (4 each, interior range), then 8 corners
    int I, J, K;
    int u = t;
    const int e0 = 4*L.N[0], e1 = 4*L.N[1], e2 = 4*L.N[2];
    if (u < e0)
    {
        const int c = u/L.N[0];
        I = 1 + (u - c*L.N[0]);
        J = (c & 1) ? L.P[1] - 1 : 0;
        K = (c & 2) ? L.P[2] - 1 : 0;
    }
    else if ((u -= e0) < e1)
    {
        const int c = u/L.N[1];
        J = 1 + (u - c*L.N[1]);
        I = (c & 1) ? L.P[0] - 1 : 0;
        K = (c & 2) ? L.P[2] - 1 : 0;
    }
    else if ((u -= e1) < e2)
    {
        const int c = u/L.N[2];
        K = 1 + (u - c*L.N[2]);
        I = (c & 1) ? L.P[0] - 1 : 0;
        J = (c & 2) ? L.P[1] - 1 : 0;
    }
    else
    {
        u -= e2;
        I = (u & 1) ? L.P[0] - 1 : 0;
        J = (u & 2) ? L.P[1] - 1 : 0;
        K = (u & 4) ? L.P[2] - 1 : 0;
    }
    edge_ghost(L, x, I, J, K);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bc_fill(LevelDims L, OpView<T> op,
                                                 Pair<T> x, int mode,
                                                 int level0, BcVals bv,
                                                 BcGroups gr, SnapPtrs<T> sn)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    bc_fill_body<T>(bid, L, op, x, mode, level0, bv, gr, sn);
}

template <typename T>
struct KBcFill
{
    static constexpr int bx = 256, by = 1;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims L;
        OpView<T> op;
        Pair<T> x;
        int mode;
        int level0;
        BcVals bv;
        BcGroups gr;
        SnapPtrs<T> sn;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        bc_fill_body<T>(bid, a.L, a.op, a.x, a.mode, a.level0, a.bv, a.gr,
                        a.sn);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_bc_fill<T>), g, dim3(bx, by), shm, s, a.L, a.op,
                           a.x, a.mode, a.level0, a.bv, a.gr, a.sn);
    }
};

// A red-black phase over the interior region and the periodic ghost fill
// with slab-plane snapshots in one launch (workgroups [0, nph) the phase,
// the rest the fill).  The fill only reads boundary-slab cells and writes
// periodic ghosts and snapshots; the interior phase writes none of these
// and reads no ghost or snapshot, so the two halves are independent and
// give the bits of the two launches in sequence.
template <typename T, int OP>
__global__ __launch_bounds__(BX*BY) void k_rbgs_fill(
    typename KRbgs<T, OP>::Args ph, typename KBcFill<T>::Args fl, unsigned nph)
{
    if (blockIdx.x < nph)
        rbgs_body<T, OP>(Bid{blockIdx.x, 0, 0}, ph.L, ph.op, ph.xc, ph.xo,
                         ph.bcl, ph.colour, ph.rg, ph.tl);
    else
        bc_fill_body<T>(Bid{blockIdx.x - nph, 0, 0}, fl.L, fl.op, fl.x,
                        fl.mode, fl.level0, fl.bv, fl.gr, fl.sn,
                        int(threadIdx.x + BX*threadIdx.y));
}



template <typename T>
__device__ __forceinline__ void fill_body(const Bid bid, T* a, T* b, long n,
                                          T v)
{
    const long t = (long)bid.x*blockDim.x + threadIdx.x;
    if (t < n) { a[t] = v; b[t] = v; }
}

template <typename T>
__global__ void k_fill(T* a, T* b, long n, T v)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    fill_body<T>(bid, a, b, n, v);
}

template <typename T>
struct KFill
{
    static constexpr int bx = 256, by = 1;
    static constexpr bool lds = false;
    struct Args
    {
        T* a;
        T* b;
        long n;
        T v;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        fill_body<T>(bid, a.a, a.b, a.n, a.v);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_fill<T>), g, dim3(bx, by), shm, s, a.a, a.b, a.n,
                           a.v);
    }
};


// ---------------------------------------------------------------------------
// Layout conversions between the reference block layout and colour-split.
template <typename T>
__device__ __forceinline__ void ref_to_split_body(const Bid bid,
                                                  const LevelDims& L,
                                                  const double* ref,
                                                  Pair<T> out)
{
    const int I = bid.z;
    const int J = bid.y*BY + threadIdx.y;
    const int p = bid.x*BX + threadIdx.x;
    if (J >= L.P[1] || p >= L.Hz) return;
    const int idx = (I*L.P[1] + J)*L.Hz + p;
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K >= L.P[2]) continue;
        col(out, (I + J + K) & 1)[idx] =
            T(ref[((long)I*L.P[1] + J)*L.P[2] + K]);
    }
}

template <typename T>
__global__ __launch_bounds__(BX*BY) void k_ref_to_split(LevelDims L,
                                                        const double* ref,
                                                        Pair<T> out)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    ref_to_split_body<T>(bid, L, ref, out);
}

template <typename T>
struct KRefToSplit
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims L;
        const double* ref;
        Pair<T> out;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        ref_to_split_body<T>(bid, a.L, a.ref, a.out);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_ref_to_split<T>), g, dim3(bx, by), shm, s, a.L,
                           a.ref, a.out);
    }
};


template <typename T>
__device__ __forceinline__ void split_to_ref_body(const Bid bid,
                                                  const LevelDims& L,
                                                  Pair<const T> in,
                                                  double* ref)
{
    const int I = bid.z;
    const int J = bid.y*BY + threadIdx.y;
    const int p = bid.x*BX + threadIdx.x;
    if (J >= L.P[1] || p >= L.Hz) return;
    const int idx = (I*L.P[1] + J)*L.Hz + p;
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K >= L.P[2]) continue;
        ref[((long)I*L.P[1] + J)*L.P[2] + K] =
            double(col(in, (I + J + K) & 1)[idx]);
    }
}

template <typename T>
__global__ __launch_bounds__(BX*BY) void k_split_to_ref(LevelDims L,
                                                        Pair<const T> in,
                                                        double* ref)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    split_to_ref_body<T>(bid, L, in, ref);
}

template <typename T>
struct KSplitToRef
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims L;
        Pair<const T> in;
        double* ref;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        split_to_ref_body<T>(bid, a.L, a.in, a.ref);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_split_to_ref<T>), g, dim3(bx, by), shm, s, a.L,
                           a.in, a.ref);
    }
};


__global__ __launch_bounds__(BX*BY) void k_stencil_ref_to_split(
    LevelDims L, const double* ref7, Pair<double> a7)
{
    const int I = blockIdx.z;
    const int J = blockIdx.y*BY + threadIdx.y;
    const int p = blockIdx.x*BX + threadIdx.x;
    if (J >= L.P[1] || p >= L.Hz) return;
    const int idx = (I*L.P[1] + J)*L.Hz + p;
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K >= L.P[2]) continue;
        const long lr = ((long)I*L.P[1] + J)*L.P[2] + K;
        double* dst = a7.c[(I + J + K) & 1];
        for (int s = 0; s < 7; s++) dst[s*L.ncol + idx] = ref7[lr*7 + s];
    }
}

__global__ __launch_bounds__(BX*BY) void k_stencil_split_to_ref(
    LevelDims L, Pair<const double> a7, double* ref7)
{
    const int I = blockIdx.z;
    const int J = blockIdx.y*BY + threadIdx.y;
    const int p = blockIdx.x*BX + threadIdx.x;
    if (J >= L.P[1] || p >= L.Hz) return;
    const int idx = (I*L.P[1] + J)*L.Hz + p;
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K >= L.P[2]) continue;
        const long lr = ((long)I*L.P[1] + J)*L.P[2] + K;
        const double* src = a7.c[(I + J + K) & 1];
        for (int s = 0; s < 7; s++) ref7[lr*7 + s] = src[s*L.ncol + idx];
    }
}

// Materialise the assembled (not yet eliminated) stencil of a
// lambda-free or face-coefficient level into the stencil representation.
template <typename T, int OP>
__global__ __launch_bounds__(BX*BY) void k_materialize(LevelDims L,
                                                       OpView<T> op,
                                                       Pair<T> a7)
{
    const int I = 1 + blockIdx.z;
    const int J = 1 + blockIdx.y*BY + threadIdx.y;
    const int p = blockIdx.x*BX + threadIdx.x;
    if (J > L.N[1]) return;
    const int row = (I*L.P[1] + J)*L.Hz;
    const int idx = row + p;
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K < 1 || K > L.N[2]) continue;
        const int colour = (I + J + K) & 1;
        T a[7];
        raw_coeffs<T, OP>(L, op, colour, idx, row, p, h, I, J, K, a);
        for (int s = 0; s < 7; s++) a7.c[colour][s*L.ncol + idx] = a[s];
    }
}

// Global cell number test of solver.C:224-228: is padded cell (I,J,K) --
// interior or a periodic face ghost -- global cell number 0?
__device__ __forceinline__ bool is_cell_zero(const LevelDims& L, int I, int J,
                                             int K)
{
    const int g[3] = {I, J, K};
    for (int d = 0; d < 3; d++)
    {
        int q = g[d] - 1;
        if (q < 0 || q >= L.N[d])
        {
            if (L.bc[2*d + (q >= L.N[d])] != PER) return false;
            q = (q + L.N[d]) % L.N[d];
        }
        if (q != 0) return false;
    }
    return true;
}

// setSingularityConstraint (solver.C:200-231): A(0,0,0) = diag(1) and every
// coupling to global cell 0 removed.
__device__ __forceinline__ void pin_body(const Bid bid, const LevelDims& L,
                                         Pair<double> a7)
{
    const int I = 1 + bid.z;
    const int J = 1 + bid.y*BY + threadIdx.y;
    const int p = bid.x*BX + threadIdx.x;
    if (J > L.N[1]) return;
    const int idx = (I*L.P[1] + J)*L.Hz + p;
    const int off[7][3] = {{0,0,0},{-1,0,0},{1,0,0},{0,-1,0},{0,1,0},
                           {0,0,-1},{0,0,1}};
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K < 1 || K > L.N[2]) continue;
        double* a = col(a7, (I + J + K) & 1);
        if (I == 1 && J == 1 && K == 1)
        {
            a[idx] = 1.0;
            for (int s = 1; s < 7; s++) a[s*L.ncol + idx] = 0.0;
            continue;
        }
        for (int s = 1; s < 7; s++)
            if (a[s*L.ncol + idx] != 0.0 &&
                is_cell_zero(L, I + off[s][0], J + off[s][1], K + off[s][2]))
                a[s*L.ncol + idx] = 0.0;
    }
}

__global__ __launch_bounds__(BX*BY) void k_pin(LevelDims L, Pair<double> a7)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    pin_body(bid, L, a7);
}

struct KPin
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims L;
        Pair<double> a7;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        pin_body(bid, a.L, a.a7);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_pin), g, dim3(bx, by), shm, s, a.L, a.a7);
    }
};


__global__ __launch_bounds__(BX*BY) void k_eliminate(LevelDims L,
                                                     Pair<double> a7)
{
    const int I = 1 + blockIdx.z;
    const int J = 1 + blockIdx.y*BY + threadIdx.y;
    const int p = blockIdx.x*BX + threadIdx.x;
    if (J > L.N[1]) return;
    const int idx = (I*L.P[1] + J)*L.Hz + p;
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K < 1 || K > L.N[2]) continue;
        double* s = a7.c[(I + J + K) & 1];
        double a[7];
        for (int c = 0; c < 7; c++) a[c] = s[c*L.ncol + idx];
        eliminate_cell(L, I, J, K, a);
        for (int c = 0; c < 7; c++) s[c*L.ncol + idx] = a[c];
    }
}

// Inhomogeneous source terms of eliminateGhosts on level 0, faces in order:
// Neumann b -= g*grad/delta, Dirichlet b -= g*2*value, g the assembled
// ghost coefficient.
template <int OP>
__device__ __forceinline__ void bc_source_body(const Bid bid,
                                               const LevelDims& L,
                                               const OpView<double>& op,
                                               Pair<double> b,
                                               const BcVals& bv)
{
    const int I = 1 + bid.z;
    const int J = 1 + bid.y*BY + threadIdx.y;
    const int p = bid.x*BX + threadIdx.x;
    if (J > L.N[1]) return;
    const int row = (I*L.P[1] + J)*L.Hz;
    const int idx = row + p;
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K < 1 || K > L.N[2]) continue;
        const int colour = (I + J + K) & 1;
        double a[7];
        raw_coeffs<double, OP>(L, op, colour, idx, row, p, h, I, J, K, a);
        const bool on[6] = {I == 1, I == L.N[0], J == 1, J == L.N[1],
                            K == 1, K == L.N[2]};
        const int pc[3] = {I, J, K};
        double v = col(b, colour)[idx];
        for (int f = 0; f < 6; f++)
        {
            if (!on[f] || L.bc[f] == PER) continue;
            const double g = bv.g[f] ? bc_value_at(L, bv.g[f], f, pc)
                                     : a[f + 1];
            const double bvf = bc_value(L, bv, f, pc);
            if (L.bc[f] == NEU)
            {
                const double delta = op.dl[f/2][pc[f/2] + (f & 1)];
                v = v - g*bvf/delta;
            }
            else
                v = v - g*2.0*bvf;
        }
        col(b, colour)[idx] = v;
    }
}

template <int OP>
__global__ __launch_bounds__(BX*BY) void k_bc_source(LevelDims L,
                                                     OpView<double> op,
                                                     Pair<double> b, BcVals bv)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    bc_source_body<OP>(bid, L, op, b, bv);
}

template <int OP>
struct KBcSource
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims L;
        OpView<double> op;
        Pair<double> b;
        BcVals bv;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        bc_source_body<OP>(bid, a.L, a.op, a.b, a.bv);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_bc_source<OP>), g, dim3(bx, by), shm, s, a.L,
                           a.op, a.b, a.bv);
    }
};


// areaWeightedFaceRestrictionScheme.C:19-56 for face direction fd: coarse
// face value = sum(a*lam)/sum(a) over its fine faces.  Reference layout.
__global__ void k_lambda_restrict(LevelDims Lf, LevelDims Lc, const double* fw0,
                                  const double* fw1, const double* fw2,
                                  const double* lamf, double* lamc, int fd)
{
    const int lim0 = Lc.N[0] + (fd == 0), lim1 = Lc.N[1] + (fd == 1),
              lim2 = Lc.N[2] + (fd == 2);
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= (long)lim0*lim1*lim2) return;
    const int i = int(t/((long)lim1*lim2));
    const int j = int((t/lim2) % lim1);
    const int k = int(t % lim2);
    double area = 0.0, value = 0.0;
    for (int ox = 0; ox < (fd == 0 ? 1 : Lc.R[0]); ox++)
    for (int oy = 0; oy < (fd == 1 ? 1 : Lc.R[1]); oy++)
    for (int oz = 0; oz < (fd == 2 ? 1 : Lc.R[2]); oz++)
    {
        const int fi = i*Lc.R[0] + ox + 1;
        const int fj = j*Lc.R[1] + oy + 1;
        const int fk = k*Lc.R[2] + oz + 1;
        const double a = fd == 0 ? fw1[fj]*fw2[fk]
                       : fd == 1 ? fw0[fi]*fw2[fk]
                       : fw0[fi]*fw1[fj];
        area = area + a;
        const double t2 = a*lamf[((long)fi*Lf.P[1] + fj)*Lf.P[2] + fk];
        value = value + t2;
    }
    lamc[((long)(i + 1)*Lc.P[1] + (j + 1))*Lc.P[2] + (k + 1)] = value/area;
}

// Face coefficient area*delta*lambda of the lower face fd of every face cell
// (linearGaussLaplacianScheme.C:63-77), stored colour-split.
__global__ void k_face_coeffs(LevelDims L, const double* w0, const double* w1,
                              const double* w2, const double* dl,
                              const double* lam, int fd, Pair<double> wf)
{
    const int lim0 = L.N[0] + (fd == 0), lim1 = L.N[1] + (fd == 1),
              lim2 = L.N[2] + (fd == 2);
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= (long)lim0*lim1*lim2) return;
    const int I = 1 + int(t/((long)lim1*lim2));
    const int J = 1 + int((t/lim2) % lim1);
    const int K = 1 + int(t % lim2);
    const double fa = fd == 0 ? w1[J]*w2[K]
                    : fd == 1 ? w0[I]*w2[K]
                    : w0[I]*w1[J];
    const int pc[3] = {I, J, K};
    const double delta = dl[pc[fd]];
    double v = fa*delta;
    v = v*lam[((long)I*L.P[1] + J)*L.P[2] + K];
    wf.c[(I + J + K) & 1][fd*L.ncol + cidx(L, I, J, K)] = v;
}

// Diagonal of im::ddt(x)/dtFrac with the Euler scheme on one level:
// ((V/dt)/dtFrac) per interior cell, V = (w0*w1)*w2 (EulerDdtScheme.C:53,
// linearSystemI.H:392-410), stored colour-split.
__global__ void k_ddt_diagonal(LevelDims L, const double* w0, const double* w1,
                               const double* w2, double dt, double frac,
                               Pair<double> dg)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= (long)L.N[0]*L.N[1]*L.N[2]) return;
    const int I = 1 + int(t/((long)L.N[1]*L.N[2]));
    const int J = 1 + int((t/L.N[2]) % L.N[1]);
    const int K = 1 + int(t % L.N[2]);
    const double v = (w0[I]*w1[J])*w2[K];
    dg.c[(I + J + K) & 1][cidx(L, I, J, K)] = (v/dt)/frac;
}

// Flux of x through the lower faces normal to axis fd, reference layout
// (PoissonSolver flux, defaultPoissonSolver.C:108-121): the face gradient
// along the outward normal of the lower face, g = (x_lower - x_cell)*delta
// (linearFaceGradientScheme.C:44-53), times lambda when given, times the
// face area (the product of the two in-face widths, the magnitude of the
// hexa face-area normal of a rectilinear cell).  Faces 1..N+1 along fd,
// interior cells along the other axes.
__global__ void k_face_flux(LevelDims L, const double* x, const double* lam,
                            const double* w0, const double* w1,
                            const double* w2, const double* dl, int fd,
                            double* flux)
{
    const int lim0 = L.N[0] + (fd == 0), lim1 = L.N[1] + (fd == 1),
              lim2 = L.N[2] + (fd == 2);
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= (long)lim0*lim1*lim2) return;
    const int I = 1 + int(t/((long)lim1*lim2));
    const int J = 1 + int((t/lim2) % lim1);
    const int K = 1 + int(t % lim2);
    const long c = ((long)I*L.P[1] + J)*L.P[2] + K;
    const long o = fd == 0 ? (long)L.P[1]*L.P[2] : (fd == 1 ? L.P[2] : 1);
    const int pc = fd == 0 ? I : (fd == 1 ? J : K);
    const double d = x[c - o] - x[c];
    const double g = d*dl[pc];
    const double fa = fd == 0 ? w1[J]*w2[K] : (fd == 1 ? w0[I]*w2[K]
                                                        : w0[I]*w1[J]);
    const double v = lam ? lam[c]*g : g;
    flux[c] = v*fa;
}

// Split correction flux of splitPoissonSolver (splitPoissonSolver.C:
// 132-135): phi = ((lambda/lambda0 - 1)*g)*A on the lower faces normal to
// fd, faces 1..N+1 along fd, with the face gradient g = (x_lower - x)*delta
// and area A of k_face_flux.
__global__ void k_split_phi(LevelDims L, const double* x, const double* lam,
                            double lam0, const double* w0, const double* w1,
                            const double* w2, const double* dl, int fd,
                            double* phi)
{
    const int lim0 = L.N[0] + (fd == 0), lim1 = L.N[1] + (fd == 1),
              lim2 = L.N[2] + (fd == 2);
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= (long)lim0*lim1*lim2) return;
    const int I = 1 + int(t/((long)lim1*lim2));
    const int J = 1 + int((t/lim2) % lim1);
    const int K = 1 + int(t % lim2);
    const long c = ((long)I*L.P[1] + J)*L.P[2] + K;
    const long o = fd == 0 ? (long)L.P[1]*L.P[2] : (fd == 1 ? L.P[2] : 1);
    const int pc = fd == 0 ? I : (fd == 1 ? J : K);
    const double d = x[c - o] - x[c];
    const double g = d*dl[pc];
    const double fa = fd == 0 ? w1[J]*w2[K] : (fd == 1 ? w0[I]*w2[K]
                                                        : w0[I]*w1[J]);
    const double r = lam[c]/lam0 - 1.0;
    phi[c] = (r*g)*fa;
}

// Modified source of splitPoissonSolver (splitPoissonSolver.C:124-137):
// bHat = b/lambda0 + div(phi), div as divergenceScheme.C:66-95 accumulates
// it -- faces by direction, each cell's lower face added and its upper face
// subtracted -- times the inverse volume 1/((dx*dy)*dz).  Interior cells.
__global__ void k_split_div(LevelDims L, const double* b, double lam0,
                            const double* phi0, const double* phi1,
                            const double* phi2, const double* w0,
                            const double* w1, const double* w2, double* bhat)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= (long)L.N[0]*L.N[1]*L.N[2]) return;
    const int I = 1 + int(t/((long)L.N[1]*L.N[2]));
    const int J = 1 + int((t/L.N[2]) % L.N[1]);
    const int K = 1 + int(t % L.N[2]);
    const long c = ((long)I*L.P[1] + J)*L.P[2] + K;
    const long oI = (long)L.P[1]*L.P[2], oJ = L.P[2];
    double acc = 0.0;
    acc = acc + phi0[c];
    acc = acc - phi0[c + oI];
    acc = acc + phi1[c];
    acc = acc - phi1[c + oJ];
    acc = acc + phi2[c];
    acc = acc - phi2[c + 1];
    const double cv = (w0[I]*w1[J])*w2[K];
    const double icv = 1.0/cv;
    const double base = b ? b[c]/lam0 : 0.0;
    bhat[c] = base + acc*icv;
}

// ---------------------------------------------------------------------------
// Coarsest-level direct solve, one workgroup (APLU.C:91-128 with rhsSource,
// linearSystemAggregation.C:336-430).  The permutation is applied first and
// the unit-lower forward substitution runs column by column; for each row
// the subtractions still happen in column order, exactly as the row-wise
// LUBacksubstitute.  The upper back substitution is row-wise in one lane.
// The factors are staged in LDS when they fit.  Then the full ghost update
// of the level (APLU.C:126).
template <typename T, bool LDS>
__device__ __forceinline__ void coarse_solve_body(const Bid bid,
                                                  const LevelDims& L,
                                                  const OpView<T>& op,
                                                  Pair<T> x,
                                                  Pair<const T> b,
                                                  const double* lug,
                                                  const int* piv, int n,
                                                  int level0, const BcVals& bv)
{
    extern __shared__ double smem[];
    double* y = smem;
    const double* lu = lug;
    const int tid = threadIdx.x, nt = blockDim.x;
    if constexpr (LDS)
    {
        double* l = smem + n;
        for (int t = tid; t < n*n; t += nt) l[t] = lug[t];
        lu = l;
    }
    const int N1 = L.N[1], N2 = L.N[2];
    const int off[7][3] = {{0,0,0},{-1,0,0},{1,0,0},{0,-1,0},{0,1,0},
                           {0,0,-1},{0,0,1}};
    for (int q = tid; q < n; q += nt)
    {
        const int I = 1 + q/(N1*N2), J = 1 + (q/N2) % N1, K = 1 + q % N2;
        const int colour = (I + J + K) & 1;
        const int idx = cidx(L, I, J, K);
        double v = double(col(b, colour)[idx]);
        for (int s = 1; s < 7; s++)
        {
            const int g[3] = {I + off[s][0], J + off[s][1], K + off[s][2]};
            bool out = false;
            for (int d = 0; d < 3; d++)
                if ((g[d] < 1 || g[d] > L.N[d]) &&
                    L.bc[2*d + (g[d] > L.N[d])] != PER)
                    out = true;
            if (out)
            {
                const double a = double(op.a7[colour][s*nci(L) + idx]);
                const double t = a*double(col(x, colour ^ 1)[cidx(L, g[0], g[1], g[2])]);
                v = v - t;
            }
        }
        y[q] = v;
    }
    __syncthreads();
    if (tid == 0)
        for (int i = 0; i < n; i++)
        {
            const int ip = piv[i];
            const double t = y[ip];
            y[ip] = y[i];
            y[i] = t;
        }
    __syncthreads();
    for (int j = 0; j < n; j++)
    {
        const double yj = y[j];
        for (int i = j + 1 + tid; i < n; i += nt)
        {
            const double t = lu[i*n + j]*yj;
            y[i] = y[i] - t;
        }
        __syncthreads();
    }
    if (tid == 0)
        for (int i = n - 1; i >= 0; i--)
        {
            const double* row = lu + i*n;
            double sum = y[i];
            for (int j = i + 1; j < n; j++)
            {
                const double t = row[j]*y[j];
                sum = sum - t;
            }
            y[i] = sum/row[i];
        }
    __syncthreads();
    for (int q = tid; q < n; q += nt)
    {
        const int I = 1 + q/(N1*N2), J = 1 + (q/N2) % N1, K = 1 + q % N2;
        col(x, (I + J + K) & 1)[cidx(L, I, J, K)] = T(y[q]);
    }
    __syncthreads();
    const Shell sh = shell_all(L);
    const int ns = sh.nA + sh.nB + sh.nC;
    for (int t = tid; t < ns; t += nt)
    {
        int I, J, K;
        if (shell_cell(L, sh, t, I, J, K))
            ghost_cell(L, op, x, 1, level0, bv, I, J, K);
    }
}

template <typename T, bool LDS>
__global__ __launch_bounds__(256) void k_coarse_solve(LevelDims L,
                                                      OpView<T> op, Pair<T> x,
                                                      Pair<const T> b,
                                                      const double* lug,
                                                      const int* piv, int n,
                                                      int level0, BcVals bv)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    coarse_solve_body<T, LDS>(bid, L, op, x, b, lug, piv, n, level0, bv);
}

template <typename T, bool LDS>
struct KCoarse
{
    static constexpr int bx = 256, by = 1;
    static constexpr bool lds = true;
    struct Args
    {
        LevelDims L;
        OpView<T> op;
        Pair<T> x;
        Pair<const T> b;
        const double* lug;
        const int* piv;
        int n;
        int level0;
        BcVals bv;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        coarse_solve_body<T, LDS>(bid, a.L, a.op, a.x, a.b, a.lug, a.piv, a.n,
                                  a.level0, a.bv);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_coarse_solve<T, LDS>), g, dim3(bx, by), shm, s,
                           a.L, a.op, a.x, a.b, a.lug, a.piv, a.n, a.level0,
                           a.bv);
    }
};


// Value of lane `lane` (a compile-time lane) in every lane.
__device__ __forceinline__ double lane_value(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(int(b), lane);
    const int hi = __builtin_amdgcn_readlane(int(b >> 32), lane);
    return __longlong_as_double((long long)(unsigned)lo |
                                ((long long)hi << 32));
}

// Coarsest-level direct solve of at most N <= 64 cells in one wavefront,
// with the same operations in the same order as k_coarse_solve (APLU.C:
// 91-128; LUBacksubstitute row by row, subtractions in increasing column
// order).  Lane i holds row i of the factors in registers (loaded from the
// column-major copy, one coalesced load per column).  The system is padded
// to N with unit rows and zero right-hand sides: every padded term is an
// exact zero subtracted after all real terms of its row, so the results do
// not change.  Forward substitution broadcasts y_j and updates the rows
// below it; the backward pass keeps, per lane, the products U(i,j)*y_j of
// the columns already solved, so row i's subtraction chain waits on no
// memory.  The row permutation is applied once (precomputed from the pivot
// sequence).  Ghost update as in k_coarse_solve.
template <typename T, int N>
__device__ __forceinline__ void coarse_solve_wave_body(const Bid bid,
                                                       const LevelDims& L,
                                                       const OpView<T>& op,
                                                       Pair<T> x,
                                                       Pair<const T> b,
                                                       const double* luc,
                                                       const int* perm, int n,
                                                       int level0,
                                                       const BcVals& bv)
{
    const int lane = threadIdx.x;
    double row[N];
#pragma unroll
    for (int j = 0; j < N; j++) row[j] = lane < N ? luc[j*N + lane] : 0.0;
    const int N1 = L.N[1], N2 = L.N[2];
    const int off[7][3] = {{0,0,0},{-1,0,0},{1,0,0},{0,-1,0},{0,1,0},
                           {0,0,-1},{0,0,1}};
    double v = 0.0;
    if (lane < n)
    {
        const int q = lane;
        const int I = 1 + q/(N1*N2), J = 1 + (q/N2) % N1, K = 1 + q % N2;
        const int colour = (I + J + K) & 1;
        const int idx = cidx(L, I, J, K);
        v = double(col(b, colour)[idx]);
        for (int s = 1; s < 7; s++)
        {
            const int g[3] = {I + off[s][0], J + off[s][1], K + off[s][2]};
            bool out = false;
            for (int d = 0; d < 3; d++)
                if ((g[d] < 1 || g[d] > L.N[d]) &&
                    L.bc[2*d + (g[d] > L.N[d])] != PER)
                    out = true;
            if (out)
            {
                const double a = double(op.a7[colour][s*nci(L) + idx]);
                const double t = a*double(col(x, colour ^ 1)[cidx(L, g[0], g[1], g[2])]);
                v = v - t;
            }
        }
    }
    double y = __shfl(v, lane < n ? perm[lane] : lane);
#pragma unroll
    for (int j = 0; j < N - 1; j++)
    {
        const double yj = lane_value(y, j);
        const double t = row[j]*yj;
        const double yn = y - t;
        y = lane > j ? yn : y;
    }
#pragma unroll
    for (int i = N - 1; i >= 0; i--)
    {
        double sum = y;
#pragma unroll
        for (int j = i + 1; j < N; j++) sum = sum - row[j];
        const double yi = lane_value(sum/row[i], i);
        y = lane == i ? yi : y;
        row[i] = row[i]*yi;
    }
    if (lane < n)
    {
        const int q = lane;
        const int I = 1 + q/(N1*N2), J = 1 + (q/N2) % N1, K = 1 + q % N2;
        col(x, (I + J + K) & 1)[cidx(L, I, J, K)] = T(y);
    }
    __syncthreads();
    const Shell sh = shell_all(L);
    const int ns = sh.nA + sh.nB + sh.nC;
    for (int t = lane; t < ns; t += 64)
    {
        int I, J, K;
        if (shell_cell(L, sh, t, I, J, K))
            ghost_cell(L, op, x, 1, level0, bv, I, J, K);
    }
}

// The coarsest-level solve of the FP32 cycles of mixed precision (no
// bitwise contract), for at most 64 cells: x = A^-1 (b - out-of-system
// couplings) with the inverse precomputed in FP64 from the LU factors
// (column-major, padded to 64 with zeros), one wavefront, lane i forming row
// i from broadcasts of the right-hand side -- 64 independent multiply-adds
// instead of the substitutions' chain of dependent ones.  Right-hand side
// and ghost update as in coarse_solve_wave_body.
template <typename T>
__device__ __forceinline__ void coarse_inv_body(const Bid, const LevelDims& L,
                                                const OpView<T>& op, Pair<T> x,
                                                Pair<const T> b,
                                                const double* inv, int n,
                                                int level0, const BcVals& bv)
{
    const int lane = threadIdx.x;
    const int N1 = L.N[1], N2 = L.N[2];
    const int off[7][3] = {{0,0,0},{-1,0,0},{1,0,0},{0,-1,0},{0,1,0},
                           {0,0,-1},{0,0,1}};
    double v = 0.0;
    if (lane < n)
    {
        const int q = lane;
        const int I = 1 + q/(N1*N2), J = 1 + (q/N2) % N1, K = 1 + q % N2;
        const int colour = (I + J + K) & 1;
        const int idx = cidx(L, I, J, K);
        v = double(col(b, colour)[idx]);
        for (int s = 1; s < 7; s++)
        {
            const int g[3] = {I + off[s][0], J + off[s][1], K + off[s][2]};
            bool out = false;
            for (int d = 0; d < 3; d++)
                if ((g[d] < 1 || g[d] > L.N[d]) &&
                    L.bc[2*d + (g[d] > L.N[d])] != PER)
                    out = true;
            if (out)
            {
                const double a = double(op.a7[colour][s*nci(L) + idx]);
                v = v - a*double(col(x, colour ^ 1)[cidx(L, g[0], g[1], g[2])]);
            }
        }
    }
    // inv is 64 x 64 with zeros beyond n, v zero beyond n: a fixed trip
    // count lets the loads of 16 columns be issued together
    double y = 0.0;
#pragma unroll 16
    for (int j = 0; j < 64; j++)
        y += inv[j*64 + lane]*__shfl(v, j, 64);
    if (lane < n)
    {
        const int q = lane;
        const int I = 1 + q/(N1*N2), J = 1 + (q/N2) % N1, K = 1 + q % N2;
        col(x, (I + J + K) & 1)[cidx(L, I, J, K)] = T(y);
    }
    __syncthreads();
    const Shell sh = shell_all(L);
    const int ns = sh.nA + sh.nB + sh.nC;
    for (int t = lane; t < ns; t += 64)
    {
        int I, J, K;
        if (shell_cell(L, sh, t, I, J, K))
            ghost_cell(L, op, x, 1, level0, bv, I, J, K);
    }
}

template <typename T>
__global__ __launch_bounds__(64) void k_coarse_inv(LevelDims L, OpView<T> op,
                                                   Pair<T> x, Pair<const T> b,
                                                   const double* inv, int n,
                                                   int level0, BcVals bv)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    coarse_inv_body<T>(bid, L, op, x, b, inv, n, level0, bv);
}

template <typename T>
struct KCoarseInv
{
    static constexpr int bx = 64, by = 1;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims L;
        OpView<T> op;
        Pair<T> x;
        Pair<const T> b;
        const double* inv;
        int n;
        int level0;
        BcVals bv;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        coarse_inv_body<T>(bid, a.L, a.op, a.x, a.b, a.inv, a.n, a.level0,
                           a.bv);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_coarse_inv<T>), g, dim3(bx, by), shm, s, a.L,
                           a.op, a.x, a.b, a.inv, a.n, a.level0, a.bv);
    }
};

template <typename T, int N>
__global__ __launch_bounds__(64) void k_coarse_solve_wave(LevelDims L,
                                                          OpView<T> op,
                                                          Pair<T> x,
                                                          Pair<const T> b,
                                                          const double* luc,
                                                          const int* perm,
                                                          int n, int level0,
                                                          BcVals bv)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    coarse_solve_wave_body<T, N>(bid, L, op, x, b, luc, perm, n, level0, bv);
}

template <typename T, int N>
struct KCoarseWave
{
    static constexpr int bx = 64, by = 1;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims L;
        OpView<T> op;
        Pair<T> x;
        Pair<const T> b;
        const double* luc;
        const int* perm;
        int n;
        int level0;
        BcVals bv;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        coarse_solve_wave_body<T, N>(bid, a.L, a.op, a.x, a.b, a.luc, a.perm,
                                     a.n, a.level0, a.bv);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_coarse_solve_wave<T, N>), g, dim3(bx, by), shm,
                           s, a.L, a.op, a.x, a.b, a.luc, a.perm, a.n,
                           a.level0, a.bv);
    }
};


// ---------------------------------------------------------------------------
// Reductions over interior cells (per-block partials, fixed order).
template <typename T, int MODE>
__device__ __forceinline__ void reduce_field_body(const Bid bid,
                                                  const LevelDims& L,
                                                  Pair<const T> f,
                                                  const Tiles& tl,
                                                  double* partials)
{
    __shared__ double sm[BX*BY/64];
    int bx, by, bz;
    if (!tile_of(tl, bid.x, bx, by, bz)) return;
    PairPos c;
    double acc = 0.0;
    if (pair_pos(L, bx, by, bz, c))
    {
        const double uA = double(ld(col(f, c.cA), c.idx));
        const double uB = double(ld(col(f, c.cB), c.idx + 1));
        acc += MODE == 0 ? uA*uA : uA;
        if (c.vB) acc += MODE == 0 ? uB*uB : uB;
    }
    const double s = block_sum(acc, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partials[tile_index(tl, bx, by, bz)] = s;
}

template <typename T, int MODE>
__global__ __launch_bounds__(BX*BY) void k_reduce_field(LevelDims L,
                                                        Pair<const T> f,
                                                        Tiles tl,
                                                        double* partials)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    reduce_field_body<T, MODE>(bid, L, f, tl, partials);
}

template <typename T, int MODE>
struct KReduceField
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = true;
    struct Args
    {
        LevelDims L;
        Pair<const T> f;
        Tiles tl;
        double* partials;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        reduce_field_body<T, MODE>(bid, a.L, a.f, a.tl, a.partials);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_reduce_field<T, MODE>), g, dim3(bx, by), shm, s,
                           a.L, a.f, a.tl, a.partials);
    }
};


// Terms of normFactors (solver.C:35-46): (|r - b + Ay| + |b - Ay|)^2 with
// Ay = rowSum(A)*y, rowSum = C + L + R + B + T + A + F in that order.
template <typename T, int OP>
__device__ __forceinline__ void nf_terms_body(const Bid bid,
                                              const LevelDims& L,
                                              const OpView<T>& op,
                                              Pair<const T> r,
                                              Pair<const T> b, double y,
                                              const Tiles& tl,
                                              double* partials)
{
    __shared__ double sm[BX*BY/64];
    int bx, by, bz;
    if (!tile_of(tl, bid.x, bx, by, bz)) return;
    PairPos c;
    double acc = 0.0;
    if (pair_pos(L, bx, by, bz, c))
    {
        for (int h = 0; h < 2; h++)
        {
            if (h == 1 && !c.vB) continue;
            const int colour = h ? c.cB : c.cA;
            const unsigned idx = c.idx + h;
            T a[7];
            coeffs<T, OP>(L, op, colour, int(idx), c.row, c.p + h, 1 - h, c.I,
                          c.J, 2*c.p + 1 + h, a);
            double rs = double(a[0]);
            for (int k = 1; k < 7; k++) rs = rs + double(a[k]);
            const double Ay = rs*y;
            const double rv = double(ld(col(r, colour), idx));
            const double bv = double(ld(col(b, colour), idx));
            const double t = fabs((rv - bv) + Ay) + fabs(bv - Ay);
            acc += t*t;
        }
    }
    const double s = block_sum(acc, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partials[tile_index(tl, bx, by, bz)] = s;
}

template <typename T, int OP>
__global__ __launch_bounds__(BX*BY) void k_nf_terms(LevelDims L, OpView<T> op,
                                                    Pair<const T> r,
                                                    Pair<const T> b, double y,
                                                    Tiles tl, double* partials)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    nf_terms_body<T, OP>(bid, L, op, r, b, y, tl, partials);
}

template <typename T, int OP>
struct KNfTerms
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = true;
    struct Args
    {
        LevelDims L;
        OpView<T> op;
        Pair<const T> r;
        Pair<const T> b;
        double y;
        Tiles tl;
        double* partials;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        nf_terms_body<T, OP>(bid, a.L, a.op, a.r, a.b, a.y, a.tl, a.partials);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_nf_terms<T, OP>), g, dim3(bx, by), shm, s, a.L,
                           a.op, a.r, a.b, a.y, a.tl, a.partials);
    }
};


// Sum of row[0, n) by one workgroup of 1024 threads, in a fixed order:
// thread t adds row[t + 1024 m] for m = k mod 16 into accumulator k, 16
// loads in flight per thread (a row of 32768 partials waits for memory
// twice instead of eight times), then a fixed tree over the 16
// accumulators and block_sum over the threads.
__device__ __forceinline__ double row_sum_1024(const double* row, int n,
                                               double* sm)
{
    constexpr int U = 16;
    double a[U];
#pragma unroll
    for (int k = 0; k < U; k++) a[k] = 0.0;
    int i = threadIdx.x;
    for (; i + (U - 1)*1024 < n; i += U*1024)
    {
        double v[U];
#pragma unroll
        for (int k = 0; k < U; k++) v[k] = row[i + k*1024];
#pragma unroll
        for (int k = 0; k < U; k++) a[k] += v[k];
    }
#pragma unroll
    for (int k = 0; k < U - 1; k++)
        if (i + k*1024 < n) a[k] += row[i + k*1024];
#pragma unroll
    for (int w = U/2; w > 0; w >>= 1)
#pragma unroll
        for (int k = 0; k < w; k++) a[k] += a[k + w];
    return block_sum(a[0], sm);
}

__device__ __forceinline__ void reduce_final_body(const Bid bid,
                                                  const double* partials,
                                                  int n, double* out,
                                                  double* hostOut)
{
    __shared__ double sm[16];
    const double s = row_sum_1024(partials, n, sm);
    if (threadIdx.x == 0)
    {
        out[0] = s;
        if (hostOut) hostOut[0] = s;
    }
}

__global__ __launch_bounds__(1024) void k_reduce_final(const double* partials,
                                                       int n, double* out,
                                                       double* hostOut)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    reduce_final_body(bid, partials, n, out, hostOut);
}

struct KReduceFinal
{
    static constexpr int bx = 1024, by = 1;
    static constexpr bool lds = true;
    struct Args
    {
        const double* partials;
        int n;
        double* out;
        double* hostOut;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        reduce_final_body(bid, a.partials, a.n, a.out, a.hostOut);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_reduce_final), g, dim3(bx, by), shm, s,
                           a.partials, a.n, a.out, a.hostOut);
    }
};





// ---------------------------------------------------------------------------
// Flexible CG iteration kernels with the scalars kept on the device.
// sc[0] = p.q, sc[1] = p.r, sc[2] = r.r, sc[3] = z.q (each the sum of one
// row of per-tile partials, reduced by k_reduce_rows).

// q = A p and the per-tile partials of p.q (row 0) and p.r (row 1).
template <int OP>
__device__ __forceinline__ void apply_dots_body(const Bid bid,
                                                const LevelDims& L,
                                                const OpView<double>& op,
                                                Pair<const double> p,
                                                Pair<double> q,
                                                Pair<const double> r,
                                                const Tiles& tl,
                                                double* partials)
{
    __shared__ double sm[BX*BY/64];
    int bx, by, bz;
    if (!tile_of(tl, bid.x, bx, by, bz)) return;
    PairPos c;
    double pq = 0.0, pr = 0.0;
    if (pair_pos(L, bx, by, bz, c))
    {
        const double rA = ld(col(r, c.cA), c.idx);
        const double rB = ld(col(r, c.cB), c.idx + 1);
        double Ax[2];
        pair_Ax<double, OP>(L, op, p, c, Ax);
        const double pA = ld(col(p, c.cA), c.idx);
        const double pB = ld(col(p, c.cB), c.idx + 1);
        st(col(q, c.cA), c.idx, Ax[0]);
        pq += pA*Ax[0];
        pr += pA*rA;
        if (c.vB)
        {
            st(col(q, c.cB), c.idx + 1, Ax[1]);
            pq += pB*Ax[1];
            pr += pB*rB;
        }
    }
    const int nb = tl.gx*tl.gy*tl.gz;
    const int blk = tile_index(tl, bx, by, bz);
    double s = block_sum(pq, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0) partials[blk] = s;
    __syncthreads();
    s = block_sum(pr, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0) partials[nb + blk] = s;
}

template <int OP>
__global__ __launch_bounds__(BX*BY) void k_apply_dots(LevelDims L,
                                                      OpView<double> op,
                                                      Pair<const double> p,
                                                      Pair<double> q,
                                                      Pair<const double> r,
                                                      Tiles tl,
                                                      double* partials)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    apply_dots_body<OP>(bid, L, op, p, q, r, tl, partials);
}

template <int OP>
struct KApplyDots
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = true;
    struct Args
    {
        LevelDims L;
        OpView<double> op;
        Pair<const double> p;
        Pair<double> q;
        Pair<const double> r;
        Tiles tl;
        double* partials;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        apply_dots_body<OP>(bid, a.L, a.op, a.p, a.q, a.r, a.tl, a.partials);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_apply_dots<OP>), g, dim3(bx, by), shm, s, a.L,
                           a.op, a.p, a.q, a.r, a.tl, a.partials);
    }
};

// r -= alpha q with alpha = sc[1]/sc[0], per-tile partials of r.r, and r
// handed to the preconditioner: STAGE 1 as FP32 into rf (mixed cycle),
// STAGE 0 as FP64 into both r0 and b0 (the level-0 right-hand side and
// restricted residual of an FP64 cycle), STAGE 2 nowhere (the learned
// preconditioner reads r itself).  x += alpha p is done by update_p_body.
template <int STAGE>
__device__ __forceinline__ void fcg_update_body(const Bid bid,
                                                const LevelDims& L,
                                                const double* sc,
                                                Pair<double> r,
                                                Pair<const double> q,
                                                Pair<float> rf,
                                                Pair<double> r0,
                                                Pair<double> b0,
                                                const Tiles& tl,
                                                double* partials)
{
    __shared__ double sm[BX*BY/64];
    int bx, by, bz;
    if (!tile_of(tl, bid.x, bx, by, bz)) return;
    const double alpha = sc[1]/sc[0];
    PairPos c;
    double acc = 0.0;
    if (pair_pos(L, bx, by, bz, c))
    {
        double rv[2], qv[2];
        for (int h = 0; h < 2; h++)
        {
            const int k = h ? c.cB : c.cA;
            const unsigned i = c.idx + h;
            rv[h] = ld(col(r, k), i);
            qv[h] = ld(col(q, k), i);
        }
        for (int h = 0; h < 2; h++)
        {
            if (h == 1 && !c.vB) continue;
            const int k = h ? c.cB : c.cA;
            const unsigned i = c.idx + h;
            const double rn = rv[h] - alpha*qv[h];
            st(col(r, k), i, rn);
            if (STAGE == 1)
                st(col(rf, k), i, float(rn));
            else if (STAGE == 0)
            {
                st(col(r0, k), i, rn);
                st(col(b0, k), i, rn);
            }
            acc += rn*rn;
        }
    }
    const double s = block_sum(acc, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partials[tile_index(tl, bx, by, bz)] = s;
}

template <int STAGE>
__global__ __launch_bounds__(BX*BY) void k_fcg_update(LevelDims L,
                                                      const double* sc,
                                                      Pair<double> r,
                                                      Pair<const double> q,
                                                      Pair<float> rf,
                                                      Pair<double> r0,
                                                      Pair<double> b0,
                                                      Tiles tl,
                                                      double* partials)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    fcg_update_body<STAGE>(bid, L, sc, r, q, rf, r0, b0, tl, partials);
}

template <int STAGE>
struct KFcgUpdate
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = true;
    struct Args
    {
        LevelDims L;
        const double* sc;
        Pair<double> r;
        Pair<const double> q;
        Pair<float> rf;
        Pair<double> r0;
        Pair<double> b0;
        Tiles tl;
        double* partials;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        fcg_update_body<STAGE>(bid, a.L, a.sc, a.r, a.q, a.rf, a.r0, a.b0,
                               a.tl, a.partials);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_fcg_update<STAGE>), g, dim3(bx, by), shm, s, a.L,
                           a.sc, a.r, a.q, a.rf, a.r0, a.b0, a.tl, a.partials);
    }
};

// per-tile partials of z.q, z in FP64 or FP32
template <typename Z>
__device__ __forceinline__ void dot_zq_body(const Bid bid, const LevelDims& L,
                                            Pair<const Z> z,
                                            Pair<const double> q,
                                            const Tiles& tl, double* partials)
{
    __shared__ double sm[BX*BY/64];
    int bx, by, bz;
    if (!tile_of(tl, bid.x, bx, by, bz)) return;
    PairPos c;
    double acc = 0.0;
    if (pair_pos(L, bx, by, bz, c))
    {
        // both cells' values loaded before any arithmetic (position idx+1
        // lies inside the padded row also where the pair has no cell B)
        double zv[2], qv[2];
        for (int h = 0; h < 2; h++)
        {
            const int k = h ? c.cB : c.cA;
            const unsigned i = c.idx + h;
            zv[h] = double(ld(col(z, k), i));
            qv[h] = ld(col(q, k), i);
        }
        acc += zv[0]*qv[0];
        const double accB = acc + zv[1]*qv[1];
        acc = c.vB ? accB : acc;
    }
    const double s = block_sum(acc, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partials[tile_index(tl, bx, by, bz)] = s;
}

template <typename Z>
__global__ __launch_bounds__(BX*BY) void k_dot_zq(LevelDims L, Pair<const Z> z,
                                                  Pair<const double> q,
                                                  Tiles tl, double* partials)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    dot_zq_body<Z>(bid, L, z, q, tl, partials);
}

template <typename Z>
struct KDotZq
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = true;
    struct Args
    {
        LevelDims L;
        Pair<const Z> z;
        Pair<const double> q;
        Tiles tl;
        double* partials;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        dot_zq_body<Z>(bid, a.L, a.z, a.q, a.tl, a.partials);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_dot_zq<Z>), g, dim3(bx, by), shm, s, a.L, a.z,
                           a.q, a.tl, a.partials);
    }
};

// p = z + beta p on interior cells, beta = -sc[3]/sc[0]
template <typename Z>
__device__ __forceinline__ void update_p_body(const Bid bid,
                                              const LevelDims& L,
                                              const double* sc,
                                              Pair<const Z> z,
                                              Pair<double> p, Pair<double> x,
                                              const Tiles& tl)
{
    int bx, by, bz;
    if (!tile_of(tl, bid.x, bx, by, bz)) return;
    const double alpha = sc[1]/sc[0];
    const double beta = -sc[3]/sc[0];
    PairPos c;
    if (!pair_pos(L, bx, by, bz, c)) return;
    // both cells' values loaded before any store (position idx+1 lies
    // inside the padded row also where the pair has no cell B)
    double zv[2], pv[2], xv[2];
    for (int h = 0; h < 2; h++)
    {
        const int k = h ? c.cB : c.cA;
        const unsigned i = c.idx + h;
        zv[h] = double(ld(col(z, k), i));
        pv[h] = ld(col(p, k), i);
        xv[h] = ld(col(x, k), i);
    }
    st(col(x, c.cA), c.idx, xv[0] + alpha*pv[0]);
    st(col(p, c.cA), c.idx, zv[0] + beta*pv[0]);
    if (c.vB)
    {
        st(col(x, c.cB), c.idx + 1, xv[1] + alpha*pv[1]);
        st(col(p, c.cB), c.idx + 1, zv[1] + beta*pv[1]);
    }
}

template <typename Z>
__global__ __launch_bounds__(BX*BY) void k_update_p(LevelDims L,
                                                    const double* sc,
                                                    Pair<const Z> z,
                                                    Pair<double> p,
                                                    Pair<double> x, Tiles tl)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    update_p_body<Z>(bid, L, sc, z, p, x, tl);
}

template <typename Z>
struct KUpdateP
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims L;
        const double* sc;
        Pair<const Z> z;
        Pair<double> p;
        Pair<double> x;
        Tiles tl;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        update_p_body<Z>(bid, a.L, a.sc, a.z, a.p, a.x, a.tl);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_update_p<Z>), g, dim3(bx, by), shm, s, a.L, a.sc,
                           a.z, a.p, a.x, a.tl);
    }
};

// Interior of a level-0 FP64 colour-split field to a dense FP32 array in the
// reference cell order (m = ((I-1)*N1 + J-1)*N2 + K-1), scaled by
// 1/rms: rms = sqrt(ss/ncells) with ss from *ssDev when given, else ssHost.
__global__ __launch_bounds__(BX*BY) void k_split_to_dense(LevelDims L,
                                                          Pair<const double> r,
                                                          const double* ssDev,
                                                          double ssHost,
                                                          float* out, Tiles tl)
{
    int bx, by, bz;
    if (!tile_of(tl, blockIdx.x, bx, by, bz)) return;
    PairPos c;
    if (!pair_pos(L, bx, by, bz, c)) return;
    const double ncell = double(L.N[0])*L.N[1]*L.N[2];
    const double ss = ssDev ? *ssDev : ssHost;
    const double inv = ss > 0.0 ? 1.0/sqrt(ss/ncell) : 0.0;
    const long base = (long(c.I - 1)*L.N[1] + (c.J - 1))*L.N[2];
    for (int h = 0; h < 2; h++)
    {
        if (h == 1 && !c.vB) continue;
        const int k = h ? c.cB : c.cA;
        const int K = 2*c.p + 1 + h;
        out[base + K - 1] = float(ld(col(r, k), c.idx + h)*inv);
    }
}

// The inverse map: z (interior) = double(in[m]) * rms / h, rms from the
// same sum of squares.
__global__ __launch_bounds__(BX*BY) void k_dense_to_split(LevelDims L,
                                                          const float* in,
                                                          const double* ssDev,
                                                          double ssHost,
                                                          double invh,
                                                          Pair<double> z,
                                                          Tiles tl)
{
    int bx, by, bz;
    if (!tile_of(tl, blockIdx.x, bx, by, bz)) return;
    PairPos c;
    if (!pair_pos(L, bx, by, bz, c)) return;
    const double ncell = double(L.N[0])*L.N[1]*L.N[2];
    const double ss = ssDev ? *ssDev : ssHost;
    const double f = sqrt(ss/ncell)*invh;
    const long base = (long(c.I - 1)*L.N[1] + (c.J - 1))*L.N[2];
    for (int h = 0; h < 2; h++)
    {
        if (h == 1 && !c.vB) continue;
        const int k = h ? c.cB : c.cA;
        const int K = 2*c.p + 1 + h;
        st(col(z, k), c.idx + h, double(in[base + K - 1])*f);
    }
}

// x += z on interior cells
__global__ __launch_bounds__(BX*BY) void k_add_interior(LevelDims L,
                                                        Pair<double> x,
                                                        Pair<const double> z,
                                                        Tiles tl)
{
    int bx, by, bz;
    if (!tile_of(tl, blockIdx.x, bx, by, bz)) return;
    PairPos c;
    if (!pair_pos(L, bx, by, bz, c)) return;
    for (int h = 0; h < 2; h++)
    {
        if (h == 1 && !c.vB) continue;
        const int k = h ? c.cB : c.cA;
        st(col(x, k), c.idx + h, ld(col(x, k), c.idx + h) +
                                  ld(col(z, k), c.idx + h));
    }
}

// Sums of nrows rows of n partials each, one workgroup per row, into
// out[row] (fixed order, as k_reduce_final).
__device__ __forceinline__ void reduce_rows_body(const Bid bid,
                                                 const double* partials,
                                                 int n, double* out,
                                                 double* hostOut)
{
    __shared__ double sm[16];
    const double s = row_sum_1024(partials + size_t(bid.x)*n, n, sm);
    if (threadIdx.x == 0)
    {
        out[bid.x] = s;
        if (hostOut) hostOut[bid.x] = s;
    }
}

__global__ __launch_bounds__(1024) void k_reduce_rows(const double* partials,
                                                      int n, double* out,
                                                      double* hostOut)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    reduce_rows_body(bid, partials, n, out, hostOut);
}

struct KReduceRows
{
    static constexpr int bx = 1024, by = 1;
    static constexpr bool lds = true;
    struct Args
    {
        const double* partials;
        int n;
        double* out;
        double* hostOut;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        reduce_rows_body(bid, a.partials, a.n, a.out, a.hostOut);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_reduce_rows), g, dim3(bx, by), shm, s,
                           a.partials, a.n, a.out, a.hostOut);
    }
};


template <typename T>
__device__ __forceinline__ void copy_body(const Bid bid, long n,
                                          Pair<const T> a, Pair<T> b)
{
    const long t = (long)bid.x*256 + threadIdx.x;
    if (t < n) { b.c[0][t] = a.c[0][t]; b.c[1][t] = a.c[1][t]; }
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy(long n, Pair<const T> a,
                                              Pair<T> b)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    copy_body<T>(bid, n, a, b);
}

template <typename T>
struct KCopy
{
    static constexpr int bx = 256, by = 1;
    static constexpr bool lds = false;
    struct Args
    {
        long n;
        Pair<const T> a;
        Pair<T> b;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        copy_body<T>(bid, a.n, a.a, a.b);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_copy<T>), g, dim3(bx, by), shm, s, a.n, a.a,
                           a.b);
    }
};

__device__ __forceinline__ void to_double_body(const Bid bid, long n,
                                               Pair<const float> a,
                                               Pair<double> b)
{
    const long t = (long)bid.x*256 + threadIdx.x;
    if (t < n) { b.c[0][t] = double(a.c[0][t]); b.c[1][t] = double(a.c[1][t]); }
}

__global__ __launch_bounds__(256) void k_to_double(long n, Pair<const float> a,
                                                   Pair<double> b)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    to_double_body(bid, n, a, b);
}

struct KToDouble
{
    static constexpr int bx = 256, by = 1;
    static constexpr bool lds = false;
    struct Args
    {
        long n;
        Pair<const float> a;
        Pair<double> b;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        to_double_body(bid, a.n, a.a, a.b);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_to_double), g, dim3(bx, by), shm, s, a.n, a.a,
                           a.b);
    }
};

// diagonalSmoother::Sweep (diagonalSmoother.C:13-69): x = b / A_C.
__global__ __launch_bounds__(BX*BY) void k_diag(LevelDims L, Pair<const double> a7,
                                                Pair<double> x,
                                                Pair<const double> b)
{
    const int I = 1 + blockIdx.z;
    const int J = 1 + blockIdx.y*BY + threadIdx.y;
    const int p = blockIdx.x*BX + threadIdx.x;
    if (J > L.N[1]) return;
    const int idx = (I*L.P[1] + J)*L.Hz + p;
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K < 1 || K > L.N[2]) continue;
        const int c = (I + J + K) & 1;
        x.c[c][idx] = b.c[c][idx]/a7.c[c][idx];
    }
}

__device__ __forceinline__ void add_correction_body(const Bid bid,
                                                    const LevelDims& L,
                                                    Pair<double> x,
    Pair<const float> e, const Tiles& tl)
{
    int bx, by, bz;
    if (!tile_of(tl, bid.x, bx, by, bz)) return;
    PairPos c;
    if (!pair_pos(L, bx, by, bz, c)) return;
    const double xA = ld(col(x, c.cA), c.idx), xB = ld(col(x, c.cB), c.idx + 1);
    const double eA = double(ld(col(e, c.cA), c.idx));
    const double eB = double(ld(col(e, c.cB), c.idx + 1));
    st(col(x, c.cA), c.idx, xA + eA);
    if (c.vB) st(col(x, c.cB), c.idx + 1, xB + eB);
}

__global__ __launch_bounds__(BX*BY) void k_add_correction(LevelDims L,
                                                          Pair<double> x,
                                                          Pair<const float> e,
                                                          Tiles tl)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    add_correction_body(bid, L, x, e, tl);
}

struct KAddCorrection
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims L;
        Pair<double> x;
        Pair<const float> e;
        Tiles tl;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        add_correction_body(bid, a.L, a.x, a.e, a.tl);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_add_correction), g, dim3(bx, by), shm, s, a.L,
                           a.x, a.e, a.tl);
    }
};


__device__ __forceinline__ void to_float_body(const Bid bid, long n,
                                              Pair<const double> in,
                                              Pair<float> out)
{
    const long t = (long)bid.x*blockDim.x + threadIdx.x;
    if (t < n)
    {
        out.c[0][t] = float(in.c[0][t]);
        out.c[1][t] = float(in.c[1][t]);
    }
}

__global__ void k_to_float(long n, Pair<const double> in, Pair<float> out)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    to_float_body(bid, n, in, out);
}

struct KToFloat
{
    static constexpr int bx = 256, by = 1;
    static constexpr bool lds = false;
    struct Args
    {
        long n;
        Pair<const double> in;
        Pair<float> out;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        to_float_body(bid, a.n, a.in, a.out);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_to_float), g, dim3(bx, by), shm, s, a.n, a.in,
                           a.out);
    }
};


template <typename T>
__device__ __forceinline__ void set_cell_body(const Bid bid, Pair<T> f,
                                              long idx, int colour, T v)
{
    col(f, colour)[idx] = v;
}

template <typename T>
__global__ void k_set_cell(Pair<T> f, long idx, int colour, T v)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    set_cell_body<T>(bid, f, idx, colour, v);
}

template <typename T>
struct KSetCell
{
    static constexpr int bx = 1, by = 1;
    static constexpr bool lds = false;
    struct Args
    {
        Pair<T> f;
        long idx;
        int colour;
        T v;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        set_cell_body<T>(bid, a.f, a.idx, a.colour, a.v);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_set_cell<T>), g, dim3(bx, by), shm, s, a.f,
                           a.idx, a.colour, a.v);
    }
};


inline dim3 grid3(const Region& rg)
{
    const int nK = rg.E[2] - rg.S[2];
    const int np = (nK + 2)/2 + 1;
    return dim3((np + BX - 1)/BX, (rg.E[1] - rg.S[1] + BY - 1)/BY,
                rg.E[0] - rg.S[0]);
}

inline dim3 grid_interior(const LevelDims& L)
{
    const int np = (L.N[2] + 2)/2 + 1;
    return dim3((np + BX - 1)/BX, (L.N[1] + BY - 1)/BY, L.N[0]);
}

inline dim3 grid_padded(const LevelDims& L)
{
    return dim3((L.Hz + BX - 1)/BX, (L.P[1] + BY - 1)/BY, L.P[0]);
}

const dim3 BLK(BX, BY, 1);

Tiles make_tiles(int gx, int gy, int gz)
{
    Tiles t;
    t.gx = gx;
    t.gy = gy;
    t.gz = gz;
    t.per = (gx*gy*gz + 7)/8;
    t.ws = 6;
    return t;
}

// Cells of a row per wavefront row for rows of n cells: the width among
// 64, 32, 16 and 8 that leaves the fewest lanes idle (ties: the wider), as
// its log2.
int row_pack_shift(int n)
{
    int best = 6, waste = ((n + 63)/64)*64 - n;
    for (int ws = 5; ws >= 3; ws--)
    {
        const int w = 1 << ws;
        const int x = ((n + w - 1)/w)*w - n;
        if (x < waste) { best = ws; waste = x; }
    }
    return best;
}

Tiles packed_tiles(int nper, int nrows, int nplanes, int ws)
{
    const int rows = BY*(64 >> ws);
    Tiles t = make_tiles((nper + (1 << ws) - 1) >> ws, (nrows + rows - 1)/rows,
                         nplanes);
    t.ws = ws;
    return t;
}

inline dim3 tiles_grid(const Tiles& t) { return dim3(8*t.per); }

inline int tiles_count(const Tiles& t) { return t.gx*t.gy*t.gz; }

// tiles of the cell-pair kernels over the interior of level L
Tiles pair_tiles(const LevelDims& L)
{
    return make_tiles(((L.N[2] + 1)/2 + BX - 1)/BX, (L.N[1] + BY - 1)/BY,
                      L.N[0]);
}

// the same with short rows packed (the bodies that read tl.ws)
Tiles pair_tiles_packed(const LevelDims& L)
{
    const int n = (L.N[2] + 1)/2;
    const int ws = row_pack_shift(n);
    const int rows = BY*(64 >> ws);
    Tiles t = make_tiles((n + (1 << ws) - 1) >> ws,
                         (L.N[1] + rows - 1)/rows, L.N[0]);
    t.ws = ws;
    return t;
}

BcVals bcvals(const BcData* v)
{
    BcVals b;
    for (int f = 0; f < 6; f++)
    {
        b.v[f] = v ? v->v[f] : 0.0;
        b.p[f] = v ? v->p[f] : nullptr;
        b.g[f] = v ? v->g[f] : nullptr;
    }
    return b;
}


// ---------------------------------------------------------------------------
// Batched forms of the kernels whose specialisations differ only in the
// operator representation (and, for the slabs, the slab axis): they take
// both from the argument struct, whose layout every specialisation shares,
// so that systems whose levels store different operator kinds share one
// launch.  BatchKind maps a kernel to the form its batched launches use.
template <typename T>
struct KRbgsAny
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = false;
    using Args = typename KRbgs<T, OP_CONST>::Args;
    __device__ static void run(const Bid bid, const Args& a)
    {
        if (a.op.kind == OP_STENCIL)
            rbgs_body<T, OP_STENCIL>(bid, a.L, a.op, a.xc, a.xo, a.bcl,
                                     a.colour, a.rg, a.tl);
        else if (a.op.kind == OP_CONST)
            rbgs_body<T, OP_CONST>(bid, a.L, a.op, a.xc, a.xo, a.bcl,
                                   a.colour, a.rg, a.tl);
        else
            rbgs_body<T, OP_FACES>(bid, a.L, a.op, a.xc, a.xo, a.bcl,
                                   a.colour, a.rg, a.tl);
    }
};

template <typename T, bool NORM>
struct KResidualAny
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = true;
    using Args = typename KResidual<T, OP_CONST, NORM>::Args;
    __device__ static void run(const Bid bid, const Args& a)
    {
        if (a.op.kind == OP_STENCIL)
            residual_body<T, OP_STENCIL, NORM>(bid, a.L, a.op, a.x, a.b, a.r,
                                               a.tl, a.partials);
        else if (a.op.kind == OP_CONST)
            residual_body<T, OP_CONST, NORM>(bid, a.L, a.op, a.x, a.b, a.r,
                                             a.tl, a.partials);
        else
            residual_body<T, OP_FACES, NORM>(bid, a.L, a.op, a.x, a.b, a.r,
                                             a.tl, a.partials);
    }
};

template <typename T>
struct KSlabsAny
{
    static constexpr int bx = 256, by = 1;
    static constexpr bool lds = true;
    using Args = typename KSlabs<T, OP_CONST, 0>::Args;
    template <int OP>
    __device__ static void axis(const Bid bid, const Args& a,
                                T (*xs)[SLAB_W + 1])
    {
        if (a.sa.axis == 0)
            rbgs_slabs_body<T, OP, 0>(bid, a.L, a.op, a.x, a.b, a.sa, xs);
        else if (a.sa.axis == 1)
            rbgs_slabs_body<T, OP, 1>(bid, a.L, a.op, a.x, a.b, a.sa, xs);
        else
            rbgs_slabs_body<T, OP, 2>(bid, a.L, a.op, a.x, a.b, a.sa, xs);
    }
    __device__ static void run(const Bid bid, const Args& a)
    {
        __shared__ T xs[SLAB_W][SLAB_W + 1];
        if (a.op.kind == OP_STENCIL) axis<OP_STENCIL>(bid, a, xs);
        else if (a.op.kind == OP_CONST) axis<OP_CONST>(bid, a, xs);
        else axis<OP_FACES>(bid, a, xs);
    }
};

// ---------------------------------------------------------------------------
// Deflation of flexible CG by indicator vectors z_k of cell sets (components
// 0..m-1; label -1: none).  A z_k is non-zero only where a cell or one of
// its six neighbours is in component k; for every interior cell i and every
// component k among its own and its neighbours' labels, the entry
// (A z_k)_i = sum over the stencil positions s whose cell is in k of
// a_s(i), the solver's own eliminated coefficients in the C,L,R,B,T,A,F
// order.  Labels are given per interior cell, (i,j,k) order, k fastest;
// a neighbour across a periodic face is the wrapped cell, across any other
// face none.  count: entries found (atomic); with write, entry e goes to
// (cell[e], comp[e], val[e]) at the position its atomic add returned (the
// host sorts them).
template <int OP>
__global__ __launch_bounds__(256) void k_defl_entries(LevelDims L,
                                                      OpView<double> op,
                                                      const int* label,
                                                      unsigned* count,
                                                      int write, unsigned cap,
                                                      unsigned* cell,
                                                      int* comp, double* val)
{
    const long n = long(L.N[0])*L.N[1]*L.N[2];
    const long t = long(blockIdx.x)*256 + threadIdx.x;
    if (t >= n) return;
    const int i = int(t/(long(L.N[1])*L.N[2]));
    const int j = int((t/L.N[2]) % L.N[1]);
    const int k = int(t % L.N[2]);
    const int g[3] = {i, j, k};
    const int off[7][3] = {{0,0,0},{-1,0,0},{1,0,0},{0,-1,0},{0,1,0},
                           {0,0,-1},{0,0,1}};
    int lab[7];
#pragma unroll
    for (int s = 0; s < 7; s++)
    {
        int q[3];
        bool out = false;
#pragma unroll
        for (int d = 0; d < 3; d++)
        {
            q[d] = g[d] + off[s][d];
            if (q[d] < 0 || q[d] >= L.N[d])
            {
                if (L.bc[2*d + (q[d] >= L.N[d])] == PER)
                    q[d] = (q[d] + L.N[d]) % L.N[d];
                else
                    out = true;
            }
        }
        lab[s] = out ? -1 : label[(long(q[0])*L.N[1] + q[1])*L.N[2] + q[2]];
    }
    bool any = false;
#pragma unroll
    for (int s = 0; s < 7; s++) any = any || lab[s] >= 0;
    if (!any) return;
    const int I = i + 1, J = j + 1, K = k + 1;
    const int colour = (I + J + K) & 1;
    const int p = K >> 1, q = K & 1;
    const int row = (I*L.P[1] + J)*L.Hz;
    const int idx = row + p;
    double a[7];
    coeffs<double, OP>(L, op, colour, idx, row, p, q, I, J, K, a);
    for (int s = 0; s < 7; s++)
    {
        const int c = lab[s];
        if (c < 0) continue;
        bool first = true;
        for (int u = 0; u < s; u++) first = first && lab[u] != c;
        if (!first) continue;
        double v = 0.0;
        for (int u = 0; u < 7; u++)
            if (lab[u] == c) v = v + a[u];
        if (v == 0.0) continue;
        const unsigned e = atomicAdd(count, 1u);
        if (write && e < cap)
        {
            cell[e] = (unsigned(colour) << 31) | unsigned(idx);
            comp[e] = c;
            val[e] = v;
        }
    }
}

// Element of a colour-split field addressed by a packed cell code (colour
// in bit 31, position below).
template <typename T>
__device__ __forceinline__ T& at_code(Pair<T> f, unsigned code)
{
    return col(f, int(code >> 31))[code & 0x7fffffffu];
}

// Chunk c of a sorted list: part[c] = sum over [beg[c], end[c]) of
// val * f(cell) (val null: the plain sum of f), in a fixed order; the
// chunks of a component are consecutive and never straddle two.
template <typename T>
__global__ __launch_bounds__(256) void k_defl_dot(const unsigned* cell,
                                                  const double* val,
                                                  const int* beg,
                                                  const int* end,
                                                  Pair<const T> f,
                                                  double* part)
{
    __shared__ double sm[4];
    const int c = blockIdx.x;
    double acc = 0.0;
    for (int e = beg[c] + int(threadIdx.x); e < end[c]; e += 256)
    {
        const double v = double(at_code(f, cell[e]));
        acc += val ? val[e]*v : v;
    }
    const double s = block_sum(acc, sm);
    if (threadIdx.x == 0) part[c] = s;
}

// f(cell) += sign * (Einv w)[comp(cell)] for every cell of the components,
// w[k] the sum of the chunk partials coff[k]..coff[k+1] (each workgroup
// sums them and forms Einv w itself; m <= 64).
template <typename T>
__global__ __launch_bounds__(256) void k_defl_correct(const unsigned* cell,
                                                      const int* comp,
                                                      long ncells,
                                                      const double* einv,
                                                      const double* part,
                                                      const int* coff, int m,
                                                      double sign, Pair<T> f)
{
    __shared__ double ws[64];
    __shared__ double cs[64];
    for (int k = int(threadIdx.x); k < m; k += 256)
    {
        double v = 0.0;
        for (int c = coff[k]; c < coff[k + 1]; c++) v += part[c];
        ws[k] = v;
    }
    __syncthreads();
    for (int k = int(threadIdx.x); k < m; k += 256)
    {
        double v = 0.0;
        for (int j = 0; j < m; j++) v += einv[k*m + j]*ws[j];
        cs[k] = sign*v;
    }
    __syncthreads();
    const long t = long(blockIdx.x)*256 + threadIdx.x;
    if (t >= ncells) return;
    T& x = at_code(f, cell[t]);
    x = T(double(x) + cs[comp[t]]);
}

// Per-block maxima of |v| over both arrays of a pair (grid-stride).
__global__ __launch_bounds__(256) void k_absmax(Pair<const double> in, long n,
                                               double* partials)
{
    __shared__ double sm[256];
    double m = 0.0;
    for (long i = long(blockIdx.x)*256 + threadIdx.x; i < 2*n;
         i += long(gridDim.x)*256)
    {
        const double v = i < n ? in.c[0][i] : in.c[1][i - n];
        m = fmax(m, fabs(v));
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (int(threadIdx.x) < s)
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(256) void k_to_half(Pair<const double> in, long n,
                                                double scale,
                                                Pair<_Float16> out)
{
    const long i = long(blockIdx.x)*256 + threadIdx.x;
    if (i >= 2*n) return;
    const int c = i >= n;
    const long j = i - c*n;
    out.c[c][j] = _Float16(in.c[c][j]*scale);
}

// Zeroing x and the first red-black phase from it in one pass: every
// element of both colour arrays is set to 0, except the cells of colour
// `colour` inside rg, which get the phase's update with all neighbours 0 --
// the same operations on the same values as a fill followed by that phase
// (k_fill, then rbgs_cell reading the zeros), so the same bits.  With
// rg.wrap, a periodic face ghost gets the value of the cell it mirrors, as
// the phase's wrap_ghosts would write it.  One thread per element, both
// colours.
template <typename T, int OP>
__device__ __forceinline__ T first_phase_cell(const LevelDims& L,
                                              const OpView<T>& op,
                                              Pair<const T> b, int colour,
                                              unsigned idx, int I, int J,
                                              int p, int q, int K)
{
    const int row = (I*L.P[1] + J)*L.Hz;
    T a[7];
    coeffs<T, OP>(L, op, colour, int(idx), row, p, q, I, J, K, a);
    const T nb[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    return gs_update<T>(a, nb, ld(col(b, colour), idx));
}

template <typename T>
__device__ __forceinline__ void zero_phase_body(const Bid bid,
                                                const LevelDims& L,
                                                const OpView<T>& op,
                                                Pair<T> x, Pair<const T> b,
                                                int colour, const Region& rg)
{
    const long n = L.ncol;
    const long t = long(bid.x)*256 + threadIdx.x;
    if (t >= 2*n) return;
    const int c = t >= n;
    const unsigned idx = unsigned(t - c*n);
    T v = T(0);
    if (c == colour)
    {
        const unsigned plane = unsigned(L.sI);
        int I = int(idx/plane);
        const unsigned rem = idx - unsigned(I)*plane;
        int J = int(rem/unsigned(L.Hz));
        int p = int(rem - unsigned(J)*unsigned(L.Hz));
        const int q = (c - I - J) & 1;
        int K = 2*p + q;
        unsigned src = idx;
        if (rg.wrap)
        {
            // a periodic face ghost takes the value of the cell it mirrors
            // (an even count along the axis keeps the colour and q)
            int out = 0;
            if ((rg.wrap & 1) && (I == 0 || I == L.N[0] + 1))
            {
                I = I == 0 ? L.N[0] : 1;
                out++;
            }
            if ((rg.wrap & 2) && (J == 0 || J == L.N[1] + 1))
            {
                J = J == 0 ? L.N[1] : 1;
                out++;
            }
            if ((rg.wrap & 4) && (K == 0 || K == L.N[2] + 1))
            {
                K = K == 0 ? L.N[2] : 1;
                p = K >> 1;
                out++;
            }
            if (out == 1)
                src = unsigned((I*L.P[1] + J)*L.Hz + p);
            else if (out > 1)
                I = -1;   // edge or corner ghost: left 0, as a face fill
        }
        if (I >= rg.S[0] && I < rg.E[0] && J >= rg.S[1] && J < rg.E[1] &&
            K >= rg.S[2] && K < rg.E[2])
        {
            if (op.kind == OP_STENCIL)
                v = first_phase_cell<T, OP_STENCIL>(L, op, b, c, src, I, J, p, q, K);
            else if (op.kind == OP_CONST)
                v = first_phase_cell<T, OP_CONST>(L, op, b, c, src, I, J, p, q, K);
            else
                v = first_phase_cell<T, OP_FACES>(L, op, b, c, src, I, J, p, q, K);
        }
    }
    st(col(x, c), idx, v);
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero_phase(LevelDims L, OpView<T> op,
                                                    Pair<T> x, Pair<const T> b,
                                                    int colour, Region rg)
{
    const Bid bid{blockIdx.x, blockIdx.y, blockIdx.z};
    zero_phase_body<T>(bid, L, op, x, b, colour, rg);
}

template <typename T>
struct KZeroPhase
{
    static constexpr int bx = 256, by = 1;
    static constexpr bool lds = false;
    struct Args
    {
        LevelDims L;
        OpView<T> op;
        Pair<T> x;
        Pair<const T> b;
        int colour;
        Region rg;
    };
    __device__ static void run(const Bid bid, const Args& a)
    {
        zero_phase_body<T>(bid, a.L, a.op, a.x, a.b, a.colour, a.rg);
    }
    static void launch(dim3 g, size_t shm, hipStream_t s, const Args& a)
    {
        hipLaunchKernelGGL((k_zero_phase<T>), g, dim3(bx, by), shm, s, a.L, a.op,
                           a.x, a.b, a.colour, a.rg);
    }
};

template <typename T>
struct KNfTermsAny
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = true;
    using Args = typename KNfTerms<T, OP_CONST>::Args;
    __device__ static void run(const Bid bid, const Args& a)
    {
        if (a.op.kind == OP_STENCIL)
            nf_terms_body<T, OP_STENCIL>(bid, a.L, a.op, a.r, a.b, a.y, a.tl,
                                         a.partials);
        else if (a.op.kind == OP_CONST)
            nf_terms_body<T, OP_CONST>(bid, a.L, a.op, a.r, a.b, a.y, a.tl,
                                       a.partials);
        else
            nf_terms_body<T, OP_FACES>(bid, a.L, a.op, a.r, a.b, a.y, a.tl,
                                       a.partials);
    }
};

struct KBcSourceAny
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = false;
    using Args = KBcSource<OP_CONST>::Args;
    __device__ static void run(const Bid bid, const Args& a)
    {
        if (a.op.kind == OP_STENCIL)
            bc_source_body<OP_STENCIL>(bid, a.L, a.op, a.b, a.bv);
        else if (a.op.kind == OP_CONST)
            bc_source_body<OP_CONST>(bid, a.L, a.op, a.b, a.bv);
        else
            bc_source_body<OP_FACES>(bid, a.L, a.op, a.b, a.bv);
    }
};

struct KApplyDotsAny
{
    static constexpr int bx = BX, by = BY;
    static constexpr bool lds = true;
    using Args = KApplyDots<OP_CONST>::Args;
    __device__ static void run(const Bid bid, const Args& a)
    {
        if (a.op.kind == OP_STENCIL)
            apply_dots_body<OP_STENCIL>(bid, a.L, a.op, a.p, a.q, a.r, a.tl,
                                        a.partials);
        else if (a.op.kind == OP_CONST)
            apply_dots_body<OP_CONST>(bid, a.L, a.op, a.p, a.q, a.r, a.tl,
                                      a.partials);
        else
            apply_dots_body<OP_FACES>(bid, a.L, a.op, a.p, a.q, a.r, a.tl,
                                      a.partials);
    }
};

template <class K>
struct BatchKind { using type = K; };
template <int OP>
struct BatchKind<KApplyDots<OP>> { using type = KApplyDotsAny; };
template <typename T, int OP>
struct BatchKind<KRbgs<T, OP>> { using type = KRbgsAny<T>; };
template <typename T, int OP, bool NORM>
struct BatchKind<KResidual<T, OP, NORM>> { using type = KResidualAny<T, NORM>; };
template <typename T, int OP, int D>
struct BatchKind<KSlabs<T, OP, D>> { using type = KSlabsAny<T>; };
template <typename T, int OP>
struct BatchKind<KNfTerms<T, OP>> { using type = KNfTermsAny<T>; };
template <int OP>
struct BatchKind<KBcSource<OP>> { using type = KBcSourceAny; };

// ---------------------------------------------------------------------------
// Batched launches.  Every batchable kernel k_x has a struct K with its
// argument struct K::Args, its workgroup shape (K::bx, K::by), its body
// K::run(workgroup, args) and K::launch (one launch of k_x).  k_many runs
// one launch for many independent systems: the systems' workgroups are
// concatenated into virtual indices (each system's share starting at a
// multiple of 8, so a workgroup keeps the XCD it would have on its own), a
// fixed grid (a multiple of 8 workgroups) walks them, and each workgroup
// runs the body on its share's argument struct.  The shares and their
// number are read from device memory, so the host can drop converged
// systems between replays of a captured graph.
template <class K>
__global__ __launch_bounds__(K::bx*K::by) void k_many(
    const typename K::Args* __restrict__ args,
    const BatchItem* __restrict__ items,
    const BatchState* __restrict__ state)
{
    const int nitems = state->nitems;
    const unsigned total = state->total;
    if (total == 0) return;
    // Virtual index W = 8p + g: the workgroups of XCD g (blockIdx & 7) split
    // the positions p into contiguous chunks, so one workgroup walks its
    // shares in order and finds each next share by a short forward step.
    const unsigned g = blockIdx.x & 7u, k = blockIdx.x >> 3;
    const unsigned per = gridDim.x >> 3;
    const unsigned P = total >> 3;
    const unsigned chunk = (P + per - 1)/per;
    const unsigned p0 = k*chunk;
    const unsigned p1 = p0 + chunk < P ? p0 + chunk : P;
    if (p0 >= p1) return;
    int lo = 0;
    {
        const unsigned W = 8u*p0 + g;
        int hi = nitems - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].start <= W) lo = mid;
            else hi = mid - 1;
        }
    }
    for (unsigned p = p0; p < p1; p++)
    {
        const unsigned W = 8u*p + g;
        while (lo + 1 < nitems && items[lo + 1].start <= W) lo++;
        const unsigned w = W - items[lo].start;
        const unsigned gx = items[lo].gx, gy = items[lo].gy;
        if (w < items[lo].n)
            K::run(Bid{w % gx, (w/gx) % gy, w/(gx*gy)}, args[items[lo].arg]);
        // the next share's body reuses the workgroup's LDS
        if constexpr (K::lds) __syncthreads();
    }
}

template <class K>
char kind_tag;

// Kinds whose tile shape the batch autotuner may change (K::retile).
template <class K, class = void>
struct HasRetile : std::false_type {};
template <class K>
struct HasRetile<K, decltype(void(K::retile(
                        std::declval<typename K::Args&>(), 0)))>
    : std::true_type {};

template <class K>
int retile_args(unsigned char* bytes, int ws, unsigned grid[3])
{
    typename K::Args a;
    std::memcpy(&a, bytes, sizeof(a));
    if (ws < 3 || ws > 6 || !K::retile(a, ws)) return 0;
    std::memcpy(bytes, &a, sizeof(a));
    grid[0] = tiles_grid(a.tl).x;
    grid[1] = 1;
    grid[2] = 1;
    return 1;
}

template <class K>
void launch_many(const BatchLaunch& b, hipStream_t s)
{
    hipLaunchKernelGGL((k_many<K>), dim3(b.grid), dim3(K::bx, K::by), b.shm,
                       s, static_cast<const typename K::Args*>(b.args),
                       b.items, b.state);
}

// Launches K, or appends the launch to the calling thread's recorder.
template <class K>
void dispatch(dim3 g, size_t shm, hipStream_t s, const typename K::Args& a)
{
    if (LaunchRecorder* r = launch_recorder())
    {
        using B = typename BatchKind<K>::type;
        static_assert(sizeof(typename B::Args) == sizeof(typename K::Args),
                      "a batched form shares the argument layout");
        LaunchRecord rec;
        rec.kind = &kind_tag<B>;
        rec.tag = r->tag;
        rec.grid[0] = g.x;
        rec.grid[1] = g.y;
        rec.grid[2] = g.z;
        rec.shm = shm;
        rec.align = alignof(typename K::Args);
        rec.args.resize(sizeof(a));
        std::memcpy(rec.args.data(), &a, sizeof(a));
        rec.many = &launch_many<B>;
        if constexpr (HasRetile<K>::value)
            rec.retile = &retile_args<K>;
        else
            rec.retile = nullptr;
        r->recs.push_back(std::move(rec));
        return;
    }
    K::launch(g, shm, s, a);
}

// Launchers of kernels without a batched form refuse to record.
void direct_only(const char* what)
{
    if (launch_recorder())
        throw std::runtime_error(std::string(what) +
                                 " has no batched form");
}

}  // namespace

LaunchRecorder*& launch_recorder()
{
    thread_local LaunchRecorder* r = nullptr;
    return r;
}

void record_tag(long tag)
{
    if (LaunchRecorder* r = launch_recorder()) r->tag = tag;
}

// ---------------------------------------------------------------------------
// launchers
int pair_tile_count(const LevelDims& L) { return tiles_count(pair_tiles(L)); }

template <typename T>
void launch_rbgs(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                 Pair<const T> b, int colour, const Region& rg, hipStream_t s)
{
    launch_rbgs_fill<T>(L, op, x, b, colour, rg, 0, nullptr, nullptr, s);
}

template <typename T>
void launch_rbgs_chain(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                       Pair<const T> b, const Region& rg, const int* cols,
                       int n, int zeroFirst, hipStream_t s)
{
    direct_only("chained red-black phases");
    if (n < 1 || n > 8) throw std::runtime_error("rbgs chain: 1 to 8 phases");
    if (rg.E[0] <= rg.S[0] || rg.E[1] <= rg.S[1] || rg.E[2] <= rg.S[2])
        return;
    RbgsChain ch{};
    ch.n = n;
    for (int k = 0; k < n; k++) ch.col[k] = cols[k];
    ch.zeroFirst = zeroFirst;
    const dim3 g(1), blk(1024);
    switch (op.kind)
    {
    case OP_STENCIL:
        hipLaunchKernelGGL((k_rbgs_chain<T, OP_STENCIL>), g, blk, 0, s, L, op,
                           x, b, rg, ch);
        break;
    case OP_CONST:
        hipLaunchKernelGGL((k_rbgs_chain<T, OP_CONST>), g, blk, 0, s, L, op,
                           x, b, rg, ch);
        break;
    default:
        hipLaunchKernelGGL((k_rbgs_chain<T, OP_FACES>), g, blk, 0, s, L, op,
                           x, b, rg, ch);
        break;
    }
}

// Tiles of the fused two-phase launch: TJ rows as large as an LDS ring of
// four planes within 64 KB allows (at most 16), CI = 9 planes, 512 threads.
FusedTiles fused_tiles(const LevelDims& L, size_t elem)
{
    FusedTiles ft{};
    const size_t perRow = 4*size_t(L.Hz)*elem;
    ft.TJ = int(std::min<size_t>(16, 65536/perRow - 2));
    ft.CI = 9;
    ft.bs = 512;
    ft.nJt = (L.P[1] + ft.TJ - 1)/ft.TJ;
    ft.nIt = (L.P[0] + ft.CI - 1)/ft.CI;
    ft.per = (ft.nJt*ft.nIt + 7)/8;
    ft.magic = unsigned(((1ull << 32) + unsigned(L.Hz) - 1)/unsigned(L.Hz));
    return ft;
}

template <typename T>
int fused_rbgs_tiles(const LevelDims& L)
{
    if (65536/(4*size_t(L.Hz)*sizeof(T)) < 4) return 0;
    const FusedTiles ft = fused_tiles(L, sizeof(T));
    return ft.nJt*ft.nIt;
}

template <typename T, int BS>
void launch_rbgs2_bs(dim3 g, size_t shm, hipStream_t s, const LevelDims& L,
                     const OpView<T>& op, T* xR, const T* xBin, T* xBout,
                     const T* bR, const T* bB, int c1, const Region& rg,
                     const FusedTiles& ft)
{
    switch (op.kind)
    {
    case OP_STENCIL:
        hipLaunchKernelGGL((k_rbgs2<T, OP_STENCIL, BS>), g, dim3(BS), shm, s,
                           L, op, xR, xBin, xBout, bR, bB, c1, rg, ft);
        break;
    case OP_CONST:
        hipLaunchKernelGGL((k_rbgs2<T, OP_CONST, BS>), g, dim3(BS), shm, s, L,
                           op, xR, xBin, xBout, bR, bB, c1, rg, ft);
        break;
    default:
        hipLaunchKernelGGL((k_rbgs2<T, OP_FACES, BS>), g, dim3(BS), shm, s, L,
                           op, xR, xBin, xBout, bR, bB, c1, rg, ft);
        break;
    }
}

template <typename T>
void launch_rbgs2(const LevelDims& L, const OpView<T>& op, T* xR,
                  const T* xBin, T* xBout, const T* bR, const T* bB, int c1,
                  const Region& rg, hipStream_t s)
{
    direct_only("fused red-black phases");
    if (fused_rbgs_tiles<T>(L) == 0)
        throw std::runtime_error("fused red-black phases: rows too long for "
                                 "the LDS ring");
    const FusedTiles ft = fused_tiles(L, sizeof(T));
    // e / Hz by the multiplier is exact for e * Hz < 2^32
    if (uint64_t(ft.TJ + 2)*uint64_t(L.Hz)*uint64_t(L.Hz) >= (1ull << 32))
        throw std::runtime_error("fused red-black phases: level too large");
    const size_t shm = 4*size_t(ft.TJ + 2)*size_t(L.Hz)*sizeof(T);
    const dim3 g(8*ft.per);
    if (ft.bs == 1024)
        launch_rbgs2_bs<T, 1024>(g, shm, s, L, op, xR, xBin, xBout, bR, bB,
                                 c1, rg, ft);
    else if (ft.bs == 512)
        launch_rbgs2_bs<T, 512>(g, shm, s, L, op, xR, xBin, xBout, bR, bB, c1,
                                rg, ft);
    else
        launch_rbgs2_bs<T, 256>(g, shm, s, L, op, xR, xBin, xBout, bR, bB, c1,
                                rg, ft);
}

// Arguments of one slab launch (one face pair, or one plane when nslab =
// 1) and its grid: one workgroup per plane when the plane with its ghost
// ring fits an LDS tile (sa.tt = 0), else tiles of 16 x 16.
template <typename T>
SlabArgs<T> slab_args(const LevelDims& L, int d, const int* planes, int nslab,
                      const int* colours, int nphase, const SlabIO<T>& io,
                      dim3* grid)
{
    SlabArgs<T> sa;
    sa.plane[0] = planes[0];
    sa.plane[1] = nslab > 1 ? planes[1] : planes[0];
    sa.nphase = nphase;
    for (int z = 0; z < 2; z++)
        for (int i = 0; i < 4; i++)
            sa.colours[z][i] = i < nphase ? colours[(nslab > 1 ? z : 0)*4 + i] : 0;
    sa.src[0] = io.src[0];
    sa.src[1] = nslab > 1 ? io.src[1] : io.src[0];
    sa.ntarget = io.ntarget;
    for (int t = 0; t < 6; t++)
    {
        sa.tptr[t] = t < io.ntarget ? io.tptr[t] : nullptr;
        sa.tdim[t] = t < io.ntarget ? io.tdim[t] : 0;
        sa.tplane[t] = t < io.ntarget ? io.tplane[t] : -1;
    }
    sa.axis = d;
    const int a1 = d == 0 ? 1 : 0;
    const int a2 = d == 2 ? 1 : 2;
    const int n1 = L.N[a1], n2 = L.N[a2];
    dim3 g(1, 1, nslab);
    if (n1 + 2 <= SLAB_W && n2 + 2 <= SLAB_W)
        sa.tt = 0;
    else
    {
        sa.tt = 16;
        g = dim3((n2 + sa.tt - 1)/sa.tt, (n1 + sa.tt - 1)/sa.tt, nslab);
    }
    *grid = g;
    return sa;
}

template <typename T>
void launch_rbgs_slabs(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                       Pair<const T> b, int d, const int* planes, int nslab,
                       const int* colours, int nphase, const SlabIO<T>& io,
                       hipStream_t s)
{
    dim3 g;
    const SlabArgs<T> sa = slab_args<T>(L, d, planes, nslab, colours, nphase,
                                        io, &g);
#define BMG_SLAB(OPK, DD)                                                     \
    dispatch<KSlabs<T, OPK, DD>>(g, 0, s, {L, op, x, b, sa})
#define BMG_SLAB_D(OPK)                                                       \
    do {                                                                      \
        if (d == 0) BMG_SLAB(OPK, 0);                                         \
        else if (d == 1) BMG_SLAB(OPK, 1);                                    \
        else BMG_SLAB(OPK, 2);                                                \
    } while (0)
    if (op.kind == OP_STENCIL) BMG_SLAB_D(OP_STENCIL);
    else if (op.kind == OP_CONST) BMG_SLAB_D(OP_CONST);
    else BMG_SLAB_D(OP_FACES);
#undef BMG_SLAB_D
#undef BMG_SLAB
}

template <typename T>
void launch_jacobi(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                   Pair<const T> b, Pair<T> tmp, const Region& rg,
                   hipStream_t s)
{
    direct_only("jacobi smoothing");
    if (rg.E[0] <= rg.S[0] || rg.E[1] <= rg.S[1] || rg.E[2] <= rg.S[2]) return;
    const dim3 g = grid3(rg);
    switch (op.kind)
    {
    case OP_STENCIL:
        hipLaunchKernelGGL((k_jacobi<T, OP_STENCIL>), g, BLK, 0, s, L, op, x,
                           b, tmp, rg);
        break;
    case OP_CONST:
        hipLaunchKernelGGL((k_jacobi<T, OP_CONST>), g, BLK, 0, s, L, op, x, b,
                           tmp, rg);
        break;
    default:
        hipLaunchKernelGGL((k_jacobi<T, OP_FACES>), g, BLK, 0, s, L, op, x, b,
                           tmp, rg);
        break;
    }
    Pair<const T> y{{tmp.c[0], tmp.c[1]}};
    hipLaunchKernelGGL((k_copy_region<T>), g, BLK, 0, s, L, y, x, rg);
}

template <typename T>
void launch_lexgs(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                  Pair<const T> b, const Region& rg, hipStream_t s)
{
    direct_only("lexicographic smoothing");
    const int M[3] = {rg.E[0] - rg.S[0], rg.E[1] - rg.S[1], rg.E[2] - rg.S[2]};
    if (M[0] <= 0 || M[1] <= 0 || M[2] <= 0) return;
    for (int d = 0; d < 3; d++)
        if (rg.S[d] < 1 || rg.E[d] > L.N[d] + 1)
            throw std::runtime_error("lexicographic sweep outside the level");
    if (M[0] > 65535)
        throw std::runtime_error("lexicographic sweep: level too long along x");
    const int ns = M[0] + M[1] + M[2] - 2;
    const int w = std::min(M[1], M[2]);
    for (int pass = 0; pass < 2; pass++)
        for (int t = 0; t < ns; t++)
        {
            const int sp = pass == 0 ? t : ns - 1 - t;
            const int ilo = std::max(0, sp - (M[1] - 1) - (M[2] - 1));
            const int ihi = std::min(M[0] - 1, sp);
            const dim3 g((w + 255)/256, ihi - ilo + 1);
            switch (op.kind)
            {
            case OP_STENCIL:
                hipLaunchKernelGGL((k_lexgs_plane<T, OP_STENCIL>), g,
                                   dim3(256), 0, s, L, op, x, b, rg, sp, ilo);
                break;
            case OP_CONST:
                hipLaunchKernelGGL((k_lexgs_plane<T, OP_CONST>), g, dim3(256),
                                   0, s, L, op, x, b, rg, sp, ilo);
                break;
            default:
                hipLaunchKernelGGL((k_lexgs_plane<T, OP_FACES>), g, dim3(256),
                                   0, s, L, op, x, b, rg, sp, ilo);
                break;
            }
        }
}

template <typename T>
void launch_residual(const LevelDims& L, const OpView<T>& op, Pair<const T> x,
                     Pair<const T> b, Pair<T> r, double* partials,
                     int* nblocks, hipStream_t s)
{
    const Tiles tl = partials ? pair_tiles(L) : pair_tiles_packed(L);
    const dim3 g = tiles_grid(tl);
    if (nblocks) *nblocks = tiles_count(tl);
#define BMG_RES(OPK, NRM)                                                     \
    dispatch<KResidual<T, OPK, NRM>>(g, 0, s, {L, op, x, b, r, tl, partials})
    if (partials)
    {
        if (op.kind == OP_STENCIL) BMG_RES(OP_STENCIL, true);
        else if (op.kind == OP_CONST) BMG_RES(OP_CONST, true);
        else BMG_RES(OP_FACES, true);
    }
    else
    {
        if (op.kind == OP_STENCIL) BMG_RES(OP_STENCIL, false);
        else if (op.kind == OP_CONST) BMG_RES(OP_CONST, false);
        else BMG_RES(OP_FACES, false);
    }
#undef BMG_RES
}

template <typename T>
void launch_restrict(const LevelDims& Lf, const LevelDims& Lc,
                     const TransferView<T>& tv, Pair<const T> rf, Pair<T> bc,
                     hipStream_t s)
{
    const Tiles tl = pair_tiles_packed(Lc);
    dispatch<KRestrict<T>>(tiles_grid(tl), 0, s, {Lf, Lc, tv, rf, bc, tl});
}

template <typename T>
void launch_prolong(const LevelDims& Lf, const LevelDims& Lc,
                    const TransferView<T>& tv, Pair<const T> xc, Pair<T> xf,
                    hipStream_t s, int skipColour)
{
    if (Lf.N[0] == 2*Lc.N[0] && Lf.N[1] == 2*Lc.N[1] && Lf.N[2] == 2*Lc.N[2])
    {
        // one thread per coarse cell (k_prolong8)
        const int ws = row_pack_shift(Lc.N[2]);
        const int rows = BY*(64 >> ws);
        Tiles tl = make_tiles((Lc.N[2] + (1 << ws) - 1) >> ws,
                              (Lc.N[1] + rows - 1)/rows, Lc.N[0]);
        tl.ws = ws;
        dispatch<KProlong8<T>>(tiles_grid(tl), 0, s,
                               {Lf, Lc, tv, xc, xf, tl, skipColour});
        return;
    }
    const Tiles tl = pair_tiles_packed(Lf);
    dispatch<KProlong<T>>(tiles_grid(tl), 0, s, {Lf, Lc, tv, xc, xf, tl});
}

// Arguments of the ghost fill (groups of threads per face pair, edges and
// corners, slab-plane snapshots); total threads in *threads.
template <typename T>
typename KBcFill<T>::Args bc_fill_args(const LevelDims& L, const OpView<T>& op,
                                       Pair<T> x, int mode, int level0,
                                       const BcData* bcval,
                                       T* const (*snap)[2], int* threads)
{
    // groups holding ghosts this mode writes: mode 0 the periodic faces,
    // mode 2 the other faces, mode 1 every ghost; then the snapshots
    auto wanted = [&](int f) {
        return mode == 1 || (mode == 0) == (L.bc[f] == PER);
    };
    BcGroups gr;
    SnapPtrs<T> sn;
    const int faceCells[3] = {2*L.N[1]*L.N[2], 2*L.N[0]*L.N[2],
                              2*L.N[0]*L.N[1]};
    const int planeCells[3] = {2*L.P[1]*L.P[2], 2*L.P[0]*L.P[2],
                               2*L.P[0]*L.P[1]};
    int total = 0;
    for (int g = 0; g < 7; g++)
    {
        int n = 0;
        if (g < 3) n = (wanted(2*g) || wanted(2*g + 1)) ? faceCells[g] : 0;
        else if (g == 3) n = mode == 1 ? 4*(L.N[0] + L.N[1] + L.N[2]) + 8 : 0;
        else
        {
            const int d = g - 4;
            const bool on = snap && snap[d][0] && snap[d][1];
            sn.p[d][0] = on ? snap[d][0] : nullptr;
            sn.p[d][1] = on ? snap[d][1] : nullptr;
            n = on ? planeCells[d] : 0;
        }
        gr.start[g] = total;
        gr.count[g] = n;
        total += (n + 255)/256*256;
    }
    gr.start[7] = total;
    *threads = total;
    return {L, op, x, mode, level0, bcvals(bcval), gr, sn};
}

template <typename T>
void launch_bc_fill_snap(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                         int mode, int level0, const BcData* bcval,
                         T* const (*snap)[2], hipStream_t s)
{
    int total = 0;
    const typename KBcFill<T>::Args a =
        bc_fill_args<T>(L, op, x, mode, level0, bcval, snap, &total);
    if (total == 0) return;
    dispatch<KBcFill<T>>(dim3(total/256), 0, s, a);
}


// Tiles of a red-black phase over region rg.
template <typename T>
Tiles rbgs_tiles(const Region& rg)
{
    // a row of the region holds at most (nK + 1)/2 cells of one colour,
    // counted from the first of them (pmin in rbgs_cell)
    const int nK = rg.E[2] - rg.S[2];
    const int ws = row_pack_shift((nK + 1)/2);
    // FP32 phases on large levels: 4 (or 2) rows per thread, so that a
    // wavefront has more loads in flight (spheres 256^3 level-0 FP32 phase
    // 0.66 -> 0.79 of HBM peak, matrix-free channel 0.37 -> 0.52, measured),
    // while the launch keeps at least 1024 (256) workgroups.  FP64 phases
    // measured no faster with more rows per thread: one.
    auto tiles_at = [&](int r)
    {
        const int rows = BY*(64 >> ws)*r;
        Tiles t = make_tiles(((nK + 1)/2 + (1 << ws) - 1) >> ws,
                             (rg.E[1] - rg.S[1] + rows - 1)/rows,
                             rg.E[0] - rg.S[0]);
        t.ws = ws;
        t.rpt = r;
        return t;
    };
    Tiles tl = tiles_at(1);
    if (std::is_same<T, float>::value)
    {
        if (tiles_count(tiles_at(4)) >= 1024)
            tl = tiles_at(4);
        else if (tiles_count(tiles_at(2)) >= 256)
            tl = tiles_at(2);
    }
    return tl;
}

template <typename T, int OP>
void launch_rbgs_op(const typename KRbgs<T, OP>::Args& a, Pair<T> x,
                    int level0, const BcData* bcval, T* const (*snap)[2],
                    hipStream_t s)
{
    const dim3 g = tiles_grid(a.tl);
    int total = 0;
    typename KBcFill<T>::Args f{};
    if (snap)
        f = bc_fill_args<T>(a.L, a.op, x, 0, level0, bcval, snap, &total);
    if (!snap || total == 0 || launch_recorder())
    {
        // recorded (batched) launches keep the phase and the fill apart
        dispatch<KRbgs<T, OP>>(g, 0, s, a);
        if (snap && total > 0) dispatch<KBcFill<T>>(dim3(total/256), 0, s, f);
        return;
    }
    hipLaunchKernelGGL((k_rbgs_fill<T, OP>), dim3(g.x + total/256),
                       dim3(BX, BY), 0, s, a, f, g.x);
}

// The phase, and with snap the periodic ghost fill (mode 0) with its
// slab-plane snapshots in the same launch.
template <typename T>
void launch_rbgs_fill(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                      Pair<const T> b, int colour, const Region& rg,
                      int level0, const BcData* bcval, T* const (*snap)[2],
                      hipStream_t s)
{
    if (rg.E[0] <= rg.S[0] || rg.E[1] <= rg.S[1] || rg.E[2] <= rg.S[2])
    {
        if (snap) launch_bc_fill_snap<T>(L, op, x, 0, level0, bcval, snap, s);
        return;
    }
    const Tiles tl = rbgs_tiles<T>(rg);
    T* xc = x.c[colour];
    const T* xo = x.c[colour ^ 1];
    const T* bc = b.c[colour];
    switch (op.kind)
    {
    case OP_STENCIL:
        launch_rbgs_op<T, OP_STENCIL>({L, op, xc, xo, bc, colour, rg, tl},
                                      x, level0, bcval, snap, s);
        break;
    case OP_CONST:
        launch_rbgs_op<T, OP_CONST>({L, op, xc, xo, bc, colour, rg, tl},
                                    x, level0, bcval, snap, s);
        break;
    default:
        launch_rbgs_op<T, OP_FACES>({L, op, xc, xo, bc, colour, rg, tl},
                                    x, level0, bcval, snap, s);
        break;
    }
}


template <typename T>
void launch_bc_fill(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                    int mode, int level0, const BcData* bcval, hipStream_t s)
{
    launch_bc_fill_snap<T>(L, op, x, mode, level0, bcval, nullptr, s);
}

template <typename T>
void launch_fill(Pair<T> x, long ncol, T v, hipStream_t s)
{
    dispatch<KFill<T>>(dim3((ncol + 255)/256), 0, s, {x.c[0], x.c[1], ncol, v});
}

template <typename T>
void launch_ref_to_split(const LevelDims& L, const double* ref, Pair<T> out,
                         hipStream_t s)
{
    dispatch<KRefToSplit<T>>(grid_padded(L), 0, s, {L, ref, out});
}

template <typename T>
void launch_split_to_ref(const LevelDims& L, Pair<const T> in, double* ref,
                         hipStream_t s)
{
    dispatch<KSplitToRef<T>>(grid_padded(L), 0, s, {L, in, ref});
}

void launch_stencil_ref_to_split(const LevelDims& L, const double* ref7,
                                 Pair<double> a7, hipStream_t s)
{
    direct_only("stencil upload");
    hipLaunchKernelGGL(k_stencil_ref_to_split, grid_padded(L), BLK, 0, s, L,
                       ref7, a7);
}

void launch_stencil_split_to_ref(const LevelDims& L, Pair<const double> a7,
                                 double* ref7, hipStream_t s)
{
    direct_only("stencil download");
    hipLaunchKernelGGL(k_stencil_split_to_ref, grid_padded(L), BLK, 0, s, L,
                       a7, ref7);
}

template <typename T>
void launch_materialize_stencil(const LevelDims& L, const OpView<T>& op,
                                Pair<T> a7, hipStream_t s)
{
    direct_only("stencil materialisation");
    if (op.kind == OP_CONST)
        hipLaunchKernelGGL((k_materialize<T, OP_CONST>), grid_interior(L), BLK,
                           0, s, L, op, a7);
    else if (op.kind == OP_FACES)
        hipLaunchKernelGGL((k_materialize<T, OP_FACES>), grid_interior(L), BLK,
                           0, s, L, op, a7);
}

void launch_pin(const LevelDims& L, Pair<double> a7, hipStream_t s)
{
    dispatch<KPin>(grid_interior(L), 0, s, {L, a7});
}

void launch_eliminate(const LevelDims& L, Pair<double> a7, hipStream_t s)
{
    direct_only("ghost elimination");
    hipLaunchKernelGGL(k_eliminate, grid_interior(L), BLK, 0, s, L, a7);
}

void launch_save_ghost_coeffs(const LevelDims& L, Pair<const double> a7,
                              int f, double* out, hipStream_t s)
{
    const int d = f/2;
    const int n = L.N[d == 0 ? 1 : 0]*L.N[d == 2 ? 1 : 2];
    hipLaunchKernelGGL(k_save_ghost_coeffs, dim3((n + 255)/256), dim3(256), 0,
                       s, L, a7, f, out);
}

void launch_bc_source(const LevelDims& L, const OpView<double>& op,
                      Pair<double> b, const BcData* bcval, hipStream_t s)
{
    if (op.kind == OP_STENCIL)
        dispatch<KBcSource<OP_STENCIL>>(grid_interior(L), 0, s,
                                        {L, op, b, bcvals(bcval)});
    else if (op.kind == OP_CONST)
        dispatch<KBcSource<OP_CONST>>(grid_interior(L), 0, s,
                                      {L, op, b, bcvals(bcval)});
    else
        dispatch<KBcSource<OP_FACES>>(grid_interior(L), 0, s,
                                      {L, op, b, bcvals(bcval)});
}

void launch_lambda_restrict(const LevelDims& Lf, const LevelDims& Lc,
                            const double* fw0, const double* fw1,
                            const double* fw2, const double* lamf,
                            double* lamc, int fd, hipStream_t s)
{
    direct_only("coefficient restriction");
    const long n = long(Lc.N[0] + (fd == 0))*(Lc.N[1] + (fd == 1))*
                   (Lc.N[2] + (fd == 2));
    hipLaunchKernelGGL(k_lambda_restrict, dim3((n + 255)/256), dim3(256), 0, s,
                       Lf, Lc, fw0, fw1, fw2, lamf, lamc, fd);
}

void launch_face_coeffs(const LevelDims& L, const double* w0, const double* w1,
                        const double* w2, const double* dl, const double* lam,
                        int fd, Pair<double> wf, hipStream_t s)
{
    direct_only("face coefficients");
    const long n = long(L.N[0] + (fd == 0))*(L.N[1] + (fd == 1))*
                   (L.N[2] + (fd == 2));
    hipLaunchKernelGGL(k_face_coeffs, dim3((n + 255)/256), dim3(256), 0, s, L,
                       w0, w1, w2, dl, lam, fd, wf);
}

void launch_ddt_diagonal(const LevelDims& L, const double* w0,
                         const double* w1, const double* w2, double delta_t,
                         double dt_frac, Pair<double> dg, hipStream_t s)
{
    direct_only("time-term diagonal");
    const long n = long(L.N[0])*L.N[1]*L.N[2];
    hipLaunchKernelGGL(k_ddt_diagonal, dim3((n + 255)/256), dim3(256), 0, s, L,
                       w0, w1, w2, delta_t, dt_frac, dg);
}

void launch_split_source(const LevelDims& L, const double* x,
                         const double* const lam[3], double lam0,
                         const double* b, const double* const w[3],
                         const double* const dl[3], double* const phi[3],
                         double* bhat, hipStream_t s)
{
    direct_only("split source");
    for (int fd = 0; fd < 3; fd++)
    {
        const long n = long(L.N[0] + (fd == 0))*(L.N[1] + (fd == 1))*
                       (L.N[2] + (fd == 2));
        hipLaunchKernelGGL(k_split_phi, dim3((n + 255)/256), dim3(256), 0, s,
                           L, x, lam[fd], lam0, w[0], w[1], w[2], dl[fd], fd,
                           phi[fd]);
    }
    const long n = long(L.N[0])*L.N[1]*L.N[2];
    hipLaunchKernelGGL(k_split_div, dim3((n + 255)/256), dim3(256), 0, s, L, b,
                       lam0, phi[0], phi[1], phi[2], w[0], w[1], w[2], bhat);
}

void launch_face_flux(const LevelDims& L, const double* x, const double* lam,
                      const double* const w[3], const double* dl, int fd,
                      double* flux, hipStream_t s)
{
    direct_only("face flux");
    const long n = long(L.N[0] + (fd == 0))*(L.N[1] + (fd == 1))*
                   (L.N[2] + (fd == 2));
    hipLaunchKernelGGL(k_face_flux, dim3((n + 255)/256), dim3(256), 0, s, L,
                       x, lam, w[0], w[1], w[2], dl, fd, flux);
}

template <typename T>
void launch_zero_phase(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                       Pair<const T> b, int colour, const Region& rg,
                       hipStream_t s)
{
    dispatch<KZeroPhase<T>>(dim3(unsigned((2*L.ncol + 255)/256)), 0, s,
                            {L, op, x, b, colour, rg});
}

template <typename T>
void launch_coarse_inv(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                       Pair<const T> b, const double* inv, int n, int level0,
                       const BcData* bcval, hipStream_t s)
{
    if (n > 64) throw std::runtime_error("coarse inverse: more than 64 cells");
    dispatch<KCoarseInv<T>>(dim3(1), 0, s,
                            {L, op, x, b, inv, n, level0, bcvals(bcval)});
}

template <typename T>
void launch_coarse_solve(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                         Pair<const T> b, const double* lu, const int* piv,
                         int n, int level0, const BcData* bcval,
                         hipStream_t s)
{
    switch (coarse_wave_width(n))
    {
    case 8:
        dispatch<KCoarseWave<T, 8>>(dim3(1), 0, s,
                                    {L, op, x, b, lu, piv, n, level0,
                                     bcvals(bcval)});
        return;
    case 16:
        dispatch<KCoarseWave<T, 16>>(dim3(1), 0, s,
                                     {L, op, x, b, lu, piv, n, level0,
                                      bcvals(bcval)});
        return;
    case 32:
        dispatch<KCoarseWave<T, 32>>(dim3(1), 0, s,
                                     {L, op, x, b, lu, piv, n, level0,
                                      bcvals(bcval)});
        return;
    case 64:
        dispatch<KCoarseWave<T, 64>>(dim3(1), 0, s,
                                     {L, op, x, b, lu, piv, n, level0,
                                      bcvals(bcval)});
        return;
    default:
        break;
    }
    if (size_t(n)*(n + 1)*sizeof(double) <= 56*1024)
        dispatch<KCoarse<T, true>>(dim3(1), size_t(n)*(n + 1)*sizeof(double), s,
                                   {L, op, x, b, lu, piv, n, level0,
                                    bcvals(bcval)});
    else
        dispatch<KCoarse<T, false>>(dim3(1), size_t(n)*sizeof(double), s,
                                    {L, op, x, b, lu, piv, n, level0,
                                     bcvals(bcval)});
}

template <typename T>
void launch_sum_sq(const LevelDims& L, Pair<const T> f, double* partials,
                   int* nblocks, hipStream_t s)
{
    const Tiles tl = pair_tiles(L);
    *nblocks = tiles_count(tl);
    dispatch<KReduceField<T, 0>>(tiles_grid(tl), 0, s, {L, f, tl, partials});
}

template <typename T>
void launch_sum(const LevelDims& L, Pair<const T> f, double* partials,
                int* nblocks, hipStream_t s)
{
    const Tiles tl = pair_tiles(L);
    *nblocks = tiles_count(tl);
    dispatch<KReduceField<T, 1>>(tiles_grid(tl), 0, s, {L, f, tl, partials});
}

template <typename T>
void launch_norm_factor_terms(const LevelDims& L, const OpView<T>& op,
                              Pair<const T> r, Pair<const T> b, double y,
                              double* partials, int* nblocks, hipStream_t s)
{
    const Tiles tl = pair_tiles(L);
    const dim3 g = tiles_grid(tl);
    *nblocks = tiles_count(tl);
    if (op.kind == OP_STENCIL)
        dispatch<KNfTerms<T, OP_STENCIL>>(g, 0, s,
                                          {L, op, r, b, y, tl, partials});
    else if (op.kind == OP_CONST)
        dispatch<KNfTerms<T, OP_CONST>>(g, 0, s,
                                        {L, op, r, b, y, tl, partials});
    else
        dispatch<KNfTerms<T, OP_FACES>>(g, 0, s,
                                        {L, op, r, b, y, tl, partials});
}

void launch_reduce_final(const double* partials, int n, double* out,
                         hipStream_t s, double* hostOut)
{
    dispatch<KReduceFinal>(dim3(1), 0, s, {partials, n, out, hostOut});
}

void launch_diag(const LevelDims& L, Pair<const double> a7, Pair<double> x,
                 Pair<const double> b, hipStream_t s)
{
    direct_only("diagonal solve");
    hipLaunchKernelGGL(k_diag, grid_interior(L), BLK, 0, s, L, a7, x, b);
}




int launch_apply_dots(const LevelDims& L, const OpView<double>& op,
                      Pair<const double> p, Pair<double> q,
                      Pair<const double> r, double* partials, hipStream_t s)
{
    const Tiles tl = pair_tiles(L);
    const dim3 g = tiles_grid(tl);
    if (op.kind == OP_STENCIL)
        dispatch<KApplyDots<OP_STENCIL>>(g, 0, s, {L, op, p, q, r, tl, partials});
    else if (op.kind == OP_CONST)
        dispatch<KApplyDots<OP_CONST>>(g, 0, s, {L, op, p, q, r, tl, partials});
    else
        dispatch<KApplyDots<OP_FACES>>(g, 0, s, {L, op, p, q, r, tl, partials});
    return tiles_count(tl);
}

int launch_fcg_update(const LevelDims& L, const double* sc, Pair<double> r,
                      Pair<const double> q, Pair<float> rf, Pair<double> r0,
                      Pair<double> b0, double* partials, hipStream_t s)
{
    const Tiles tl = pair_tiles(L);
    const dim3 g = tiles_grid(tl);
    if (rf.c[0])
        dispatch<KFcgUpdate<1>>(g, 0, s, {L, sc, r, q, rf, r0, b0, tl, partials});
    else if (r0.c[0])
        dispatch<KFcgUpdate<0>>(g, 0, s, {L, sc, r, q, rf, r0, b0, tl, partials});
    else
        dispatch<KFcgUpdate<2>>(g, 0, s, {L, sc, r, q, rf, r0, b0, tl, partials});
    return tiles_count(tl);
}

void launch_split_to_dense(const LevelDims& L, Pair<const double> r,
                           const double* ssDev, double ssHost, float* out,
                           hipStream_t s)
{
    direct_only("the learned preconditioner");
    const Tiles tl = pair_tiles(L);
    hipLaunchKernelGGL(k_split_to_dense, tiles_grid(tl), BLK, 0, s, L, r,
                       ssDev, ssHost, out, tl);
}

void launch_dense_to_split(const LevelDims& L, const float* in,
                           const double* ssDev, double ssHost, double invh,
                           Pair<double> z, hipStream_t s)
{
    direct_only("the learned preconditioner");
    const Tiles tl = pair_tiles(L);
    hipLaunchKernelGGL(k_dense_to_split, tiles_grid(tl), BLK, 0, s, L, in,
                       ssDev, ssHost, invh, z, tl);
}

void launch_add_interior(const LevelDims& L, Pair<double> x,
                         Pair<const double> z, hipStream_t s)
{
    direct_only("the learned preconditioner");
    const Tiles tl = pair_tiles(L);
    hipLaunchKernelGGL(k_add_interior, tiles_grid(tl), BLK, 0, s, L, x, z, tl);
}

template <typename Z>
int launch_dot_zq(const LevelDims& L, Pair<const Z> z, Pair<const double> q,
                  double* partials, hipStream_t s)
{
    const Tiles tl = pair_tiles(L);
    dispatch<KDotZq<Z>>(tiles_grid(tl), 0, s, {L, z, q, tl, partials});
    return tiles_count(tl);
}
template int launch_dot_zq<double>(const LevelDims&, Pair<const double>,
                                   Pair<const double>, double*, hipStream_t);
template int launch_dot_zq<float>(const LevelDims&, Pair<const float>,
                                  Pair<const double>, double*, hipStream_t);

template <typename Z>
void launch_update_p(const LevelDims& L, const double* sc, Pair<const Z> z,
                     Pair<double> p, Pair<double> x, hipStream_t s)
{
    const Tiles tl = pair_tiles(L);
    dispatch<KUpdateP<Z>>(tiles_grid(tl), 0, s, {L, sc, z, p, x, tl});
}
template void launch_update_p<double>(const LevelDims&, const double*,
                                      Pair<const double>, Pair<double>,
                                      Pair<double>, hipStream_t);
template void launch_update_p<float>(const LevelDims&, const double*,
                                     Pair<const float>, Pair<double>,
                                     Pair<double>, hipStream_t);

void launch_reduce_rows(const double* partials, int n, int nrows, double* out,
                        hipStream_t s, double* hostOut)
{
    dispatch<KReduceRows>(dim3(nrows), 0, s, {partials, n, out, hostOut});
}


template <typename T>
void launch_copy(long n, Pair<const T> a, Pair<T> b, hipStream_t s)
{
    dispatch<KCopy<T>>(dim3(unsigned((n + 255)/256)), 0, s, {n, a, b});
}
template void launch_copy<double>(long, Pair<const double>, Pair<double>,
                                  hipStream_t);

void launch_to_double(long n, Pair<const float> a, Pair<double> b,
                      hipStream_t s)
{
    dispatch<KToDouble>(dim3(unsigned((n + 255)/256)), 0, s, {n, a, b});
}

void launch_add_correction(const LevelDims& L, Pair<double> x,
                           Pair<const float> e, hipStream_t s)
{
    const Tiles tl = pair_tiles(L);
    dispatch<KAddCorrection>(tiles_grid(tl), 0, s, {L, x, e, tl});
}

double device_absmax(long n, Pair<const double> in, double* partials,
                     size_t npartials, hipStream_t s)
{
    const long total = 2*n;
    int nb = int(std::min<long>((total + 255)/256, long(npartials)));
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_absmax, dim3(nb), dim3(256), 0, s, in, n, partials);
    std::vector<double> h(nb);
    if (hipMemcpyAsync(h.data(), partials, nb*sizeof(double),
                       hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        throw std::runtime_error("absmax: copy failed");
    double m = 0.0;
    for (double v : h) m = std::max(m, v);
    return m;
}

void launch_to_half(long n, Pair<const double> in, double scale,
                    Pair<_Float16> out, hipStream_t s)
{
    direct_only("half conversion");
    hipLaunchKernelGGL(k_to_half, dim3(unsigned((2*n + 255)/256)), dim3(256), 0,
                       s, in, n, scale, out);
}

void launch_to_float(long ncol, Pair<const double> in, Pair<float> out,
                     hipStream_t s)
{
    dispatch<KToFloat>(dim3((ncol + 255)/256), 0, s, {ncol, in, out});
}

unsigned launch_defl_entries(const LevelDims& L, const OpView<double>& op,
                             const int* label, unsigned* count, unsigned cap,
                             unsigned* cell, int* comp, double* val,
                             hipStream_t s)
{
    direct_only("deflation set-up");
    const long n = long(L.N[0])*L.N[1]*L.N[2];
    const dim3 g(unsigned((n + 255)/256));
    const int write = cell != nullptr;
    switch (op.kind)
    {
    case OP_STENCIL:
        hipLaunchKernelGGL((k_defl_entries<OP_STENCIL>), g, dim3(256), 0, s, L,
                           op, label, count, write, cap, cell, comp, val);
        break;
    case OP_CONST:
        hipLaunchKernelGGL((k_defl_entries<OP_CONST>), g, dim3(256), 0, s, L,
                           op, label, count, write, cap, cell, comp, val);
        break;
    default:
        hipLaunchKernelGGL((k_defl_entries<OP_FACES>), g, dim3(256), 0, s, L,
                           op, label, count, write, cap, cell, comp, val);
        break;
    }
    return 0;
}

template <typename T>
void launch_defl_dot(const unsigned* cell, const double* val, const int* beg,
                     const int* end, int nchunks, Pair<const T> f,
                     double* part, hipStream_t s)
{
    direct_only("deflation");
    if (nchunks <= 0) return;
    hipLaunchKernelGGL((k_defl_dot<T>), dim3(nchunks), dim3(256), 0, s, cell,
                       val, beg, end, f, part);
}

template <typename T>
void launch_defl_correct(const unsigned* cell, const int* comp, long ncells,
                         const double* einv, const double* part,
                         const int* coff, int m, double sign, Pair<T> f,
                         hipStream_t s)
{
    direct_only("deflation");
    if (ncells <= 0) return;
    hipLaunchKernelGGL((k_defl_correct<T>), dim3(unsigned((ncells + 255)/256)),
                       dim3(256), 0, s, cell, comp, ncells, einv, part, coff,
                       m, sign, f);
}

template <typename T>
void launch_set_cell(Pair<T> f, long idx, int colour, T v, hipStream_t s)
{
    dispatch<KSetCell<T>>(dim3(1), 0, s, {f, idx, colour, v});
}

#define BMG_INST(T)                                                           \
    template void launch_rbgs<T>(const LevelDims&, const OpView<T>&, Pair<T>, \
                                 Pair<const T>, int, const Region&,           \
                                 hipStream_t);                                \
    template void launch_rbgs_chain<T>(const LevelDims&, const OpView<T>&,    \
                                       Pair<T>, Pair<const T>, const Region&, \
                                       const int*, int, int, hipStream_t);    \
    template void launch_rbgs_fill<T>(const LevelDims&, const OpView<T>&,     \
                                      Pair<T>, Pair<const T>, int,            \
                                      const Region&, int, const BcData*,      \
                                      T* const (*)[2], hipStream_t);          \
    template void launch_rbgs2<T>(const LevelDims&, const OpView<T>&, T*,     \
                                  const T*, T*, const T*, const T*, int,      \
                                  const Region&, hipStream_t);                \
    template int fused_rbgs_tiles<T>(const LevelDims&);                        \
    template void launch_rbgs_slabs<T>(const LevelDims&, const OpView<T>&,    \
                                       Pair<T>, Pair<const T>, int,           \
                                       const int*, int, const int*, int,      \
                                       const SlabIO<T>&, hipStream_t);                          \
    template void launch_jacobi<T>(const LevelDims&, const OpView<T>&,        \
                                   Pair<T>, Pair<const T>, Pair<T>,           \
                                   const Region&, hipStream_t);               \
    template void launch_lexgs<T>(const LevelDims&, const OpView<T>&,         \
                                  Pair<T>, Pair<const T>, const Region&,      \
                                  hipStream_t);                               \
    template void launch_residual<T>(const LevelDims&, const OpView<T>&,      \
                                     Pair<const T>, Pair<const T>, Pair<T>,   \
                                     double*, int*, hipStream_t);             \
    template void launch_restrict<T>(const LevelDims&, const LevelDims&,      \
                                     const TransferView<T>&, Pair<const T>,   \
                                     Pair<T>, hipStream_t);                   \
    template void launch_prolong<T>(const LevelDims&, const LevelDims&,       \
                                    const TransferView<T>&, Pair<const T>,    \
                                    Pair<T>, hipStream_t, int);               \
    template void launch_bc_fill<T>(const LevelDims&, const OpView<T>&,       \
                                    Pair<T>, int, int, const BcData*,         \
                                    hipStream_t);                             \
    template void launch_bc_fill_snap<T>(const LevelDims&, const OpView<T>&,  \
                                         Pair<T>, int, int, const BcData*,    \
                                         T* const (*)[2], hipStream_t);       \
    template void launch_fill<T>(Pair<T>, long, T, hipStream_t);              \
    template void launch_zero_phase<T>(const LevelDims&, const OpView<T>&,    \
                                       Pair<T>, Pair<const T>, int,           \
                                       const Region&, hipStream_t);           \
    template void launch_coarse_inv<T>(const LevelDims&, const OpView<T>&,    \
                                       Pair<T>, Pair<const T>, const double*, \
                                       int, int, const BcData*, hipStream_t); \
    template void launch_defl_dot<T>(const unsigned*, const double*,          \
                                     const int*, const int*, int,             \
                                     Pair<const T>, double*, hipStream_t);    \
    template void launch_defl_correct<T>(const unsigned*, const int*, long,   \
                                         const double*, const double*,        \
                                         const int*, int, double, Pair<T>,    \
                                         hipStream_t);                        \
    template void launch_ref_to_split<T>(const LevelDims&, const double*,     \
                                         Pair<T>, hipStream_t);               \
    template void launch_split_to_ref<T>(const LevelDims&, Pair<const T>,     \
                                         double*, hipStream_t);               \
    template void launch_materialize_stencil<T>(const LevelDims&,             \
                                                const OpView<T>&, Pair<T>,    \
                                                hipStream_t);                 \
    template void launch_coarse_solve<T>(const LevelDims&, const OpView<T>&,  \
                                         Pair<T>, Pair<const T>,              \
                                         const double*, const int*, int, int, \
                                         const BcData*, hipStream_t);         \
    template void launch_sum_sq<T>(const LevelDims&, Pair<const T>, double*,  \
                                   int*, hipStream_t);                        \
    template void launch_sum<T>(const LevelDims&, Pair<const T>, double*,     \
                                int*, hipStream_t);                           \
    template void launch_norm_factor_terms<T>(const LevelDims&,               \
                                              const OpView<T>&,               \
                                              Pair<const T>, Pair<const T>,   \
                                              double, double*, int*,          \
                                              hipStream_t);                   \
    template void launch_set_cell<T>(Pair<T>, long, int, T, hipStream_t);

BMG_INST(double)
BMG_INST(float)

}  // namespace bmg
