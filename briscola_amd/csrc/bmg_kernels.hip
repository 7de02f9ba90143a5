// HIP kernels of the MI355X multigrid solver (gfx950).
//
// All kernels work on the colour-split layout described in bmg_internal.h.
// They are built with -ffp-contract=off: every product and sum is rounded in
// the order the reference writes it, so the FP64 path reproduces the
// reference arithmetic bit for bit (the order of each expression is cited at
// the kernel).
#include "bmg_internal.h"

#include <algorithm>

namespace bmg {

namespace {

constexpr int BX = 64;   // one wavefront along the contiguous axis
constexpr int BY = 4;

struct BcVals {
    double v[6];
};

__device__ __forceinline__ int cidx(const LevelDims& L, int I, int J, int K)
{
    return (I*L.P[1] + J)*L.Hz + (K >> 1);
}

__device__ __forceinline__ int nci(const LevelDims& L) { return int(L.ncol); }

// Stencil coefficients [C,L,R,B,T,A,F] of padded cell (I,J,K), as assembled
// by linearGaussLaplacianScheme.C:61-94 (centre = 0 - v_xl - v_xu - v_yl -
// v_yu - v_zl - v_zu, faces in loop order), before ghost elimination.
template <typename T, int OP>
__device__ __forceinline__ void raw_coeffs(const LevelDims& L,
                                           const OpView<T>& op, int colour,
                                           int idx, int row, int p, int q,
                                           int I, int J, int K, T a[7])
{
    if constexpr (OP == OP_STENCIL)
    {
        const T* s = op.a7[colour];
        const int n = nci(L);
#pragma unroll
        for (int c = 0; c < 7; c++) a[c] = s[c*n + idx];
    }
    else
    {
        if constexpr (OP == OP_CONST)
        {
            const T wx = op.w[0][I], wy = op.w[1][J], wz = op.w[2][K];
            const T ax = wy*wz;
            const T ay = wx*wz;
            const T az = wx*wy;
            a[1] = ax*op.dl[0][I];
            a[2] = ax*op.dl[0][I+1];
            a[3] = ay*op.dl[1][J];
            a[4] = ay*op.dl[1][J+1];
            a[5] = az*op.dl[2][K];
            a[6] = az*op.dl[2][K+1];
        }
        else
        {
            const int n = nci(L);
            const T* wc = op.wf[colour];
            const T* wo = op.wf[colour ^ 1];
            a[1] = wc[idx];
            a[2] = wo[idx + L.sI];
            a[3] = wc[n + idx];
            a[4] = wo[n + idx + L.Hz];
            a[5] = wc[2*n + idx];
            a[6] = wo[2*n + row + p + q];
        }
        T c0 = T(0) - a[1];
        c0 = c0 - a[2];
        c0 = c0 - a[3];
        c0 = c0 - a[4];
        c0 = c0 - a[5];
        c0 = c0 - a[6];
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

// ---------------------------------------------------------------------------
// Red-black Gauss-Seidel phase: update every cell of one colour inside
// region [S,E) (rbgsSmoother.C:65-183):
//   x_c = (b_c - (A1 x_-i + A2 x_+i + A3 x_-j + A4 x_+j + A5 x_-k + A6 x_+k))
//         / A0
// One thread per cell of the colour; lanes run along the contiguous axis.
template <typename T, int OP>
__global__ __launch_bounds__(BX*BY) void k_rbgs(LevelDims L, OpView<T> op,
                                                T* __restrict__ xc,
                                                const T* __restrict__ xo,
                                                const T* __restrict__ bcl,
                                                int colour, Region rg)
{
    const int I = rg.S[0] + blockIdx.z;
    const int J = rg.S[1] + blockIdx.y*BY + threadIdx.y;
    if (J >= rg.E[1]) return;
    const int q = (colour - I - J) & 1;
    const int pmin = (rg.S[2] - q + 1) >> 1;
    const int p = pmin + blockIdx.x*BX + threadIdx.x;
    const int K = 2*p + q;
    if (K >= rg.E[2]) return;
    const int sI = L.sI;
    const int row = (I*L.P[1] + J)*L.Hz;
    const int idx = row + p;
    T a[7];
    coeffs<T, OP>(L, op, colour, idx, row, p, q, I, J, K, a);
    const T P =
        a[1]*xo[idx - sI]
      + a[2]*xo[idx + sI]
      + a[3]*xo[idx - L.Hz]
      + a[4]*xo[idx + L.Hz]
      + a[5]*xo[row + p - 1 + q]
      + a[6]*xo[row + p + q];
    xc[idx] = (bcl[idx] - P)/a[0];
}

// All red-black phases of the width-1 boundary slabs of one face pair
// (rbgsSmoother::SweepBoundary, rbgsSmoother.C:186-374) in one launch.  A
// slab is a plane of cells whose out-of-plane neighbours (the interior layer
// and the ghost layer filled just before) stay fixed while its phases run,
// so every phase is a 2-D red-black update of the plane.  Each workgroup
// owns a TT x TT tile, loads the tile grown by H cells (H = number of
// phases) into LDS and runs the phases there: phase p updates the tile grown
// by H-1-p cells, whose neighbours are exact after phase p-1, so the tile is
// exact after the last phase and the halo work is redundant but identical.
// Opposite slabs of a pair are independent (they are at least two cells
// apart) and run as blockIdx.z = 0/1.  The update is the row product in the
// reference order with the two out-of-plane terms read from global memory.
struct SlabArgs {
    int d;            // slab normal axis
    int plane[2];     // padded index of each slab plane
    int nphase;       // H
    int colours[2][4];
    int tt;           // output tile edge
};

template <typename T, int OP>
__global__ __launch_bounds__(256) void k_rbgs_slabs(LevelDims L, OpView<T> op,
                                                    Pair<T> x, Pair<const T> b,
                                                    SlabArgs sa)
{
    extern __shared__ unsigned char slab_smem[];
    T* xs = reinterpret_cast<T*>(slab_smem);
    const int d = sa.d;
    const int a1 = d == 0 ? 1 : 0;
    const int a2 = d == 2 ? 1 : 2;
    const int H = sa.nphase;
    const int W = sa.tt + 2*H;
    const int Ip = sa.plane[blockIdx.z];
    const int o1 = 1 + blockIdx.y*sa.tt - H;
    const int o2 = 1 + blockIdx.x*sa.tt - H;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int e = tid; e < W*W; e += nt)
    {
        const int c1 = o1 + e/W, c2 = o2 + e % W;
        T v = T(0);
        if (c1 >= 0 && c1 < L.P[a1] && c2 >= 0 && c2 < L.P[a2])
        {
            int g[3];
            g[d] = Ip; g[a1] = c1; g[a2] = c2;
            v = x.c[(g[0] + g[1] + g[2]) & 1][cidx(L, g[0], g[1], g[2])];
        }
        xs[e] = v;
    }
    __syncthreads();
    for (int ph = 0; ph < H; ph++)
    {
        const int colour = sa.colours[blockIdx.z][ph];
        const int grow = H - 1 - ph;
        const int lo = H - grow, hi = W - (H - grow);   // LDS window
        for (int e = tid; e < W*W; e += nt)
        {
            const int u1 = e/W, u2 = e % W;
            if (u1 < lo || u1 >= hi || u2 < lo || u2 >= hi) continue;
            const int c1 = o1 + u1, c2 = o2 + u2;
            if (c1 < 1 || c1 > L.N[a1] || c2 < 1 || c2 > L.N[a2]) continue;
            int g[3];
            g[d] = Ip; g[a1] = c1; g[a2] = c2;
            const int I = g[0], J = g[1], K = g[2];
            if (((I + J + K) & 1) != colour) continue;
            const int q = K & 1;
            const int p = K >> 1;
            const int row = (I*L.P[1] + J)*L.Hz;
            const int idx = row + p;
            T a[7];
            coeffs<T, OP>(L, op, colour, idx, row, p, q, I, J, K, a);
            const T* xo = x.c[colour ^ 1];
            // neighbour values in the reference order -i,+i,-j,+j,-k,+k
            T nb[6];
            const int e1 = (u1)*W, e1m = (u1 - 1)*W, e1p = (u1 + 1)*W;
            if (d == 0)
            {
                nb[0] = xo[idx - L.sI];
                nb[1] = xo[idx + L.sI];
                nb[2] = xs[e1m + u2];
                nb[3] = xs[e1p + u2];
                nb[4] = xs[e1 + u2 - 1];
                nb[5] = xs[e1 + u2 + 1];
            }
            else if (d == 1)
            {
                nb[0] = xs[e1m + u2];
                nb[1] = xs[e1p + u2];
                nb[2] = xo[idx - L.Hz];
                nb[3] = xo[idx + L.Hz];
                nb[4] = xs[e1 + u2 - 1];
                nb[5] = xs[e1 + u2 + 1];
            }
            else
            {
                nb[0] = xs[e1m + u2];
                nb[1] = xs[e1p + u2];
                nb[2] = xs[e1 + u2 - 1];
                nb[3] = xs[e1 + u2 + 1];
                nb[4] = xo[row + p - 1 + q];
                nb[5] = xo[row + p + q];
            }
            const T P =
                a[1]*nb[0]
              + a[2]*nb[1]
              + a[3]*nb[2]
              + a[4]*nb[3]
              + a[5]*nb[4]
              + a[6]*nb[5];
            xs[e] = (b.c[colour][idx] - P)/a[0];
        }
        __syncthreads();
    }
    for (int e = tid; e < W*W; e += nt)
    {
        const int u1 = e/W, u2 = e % W;
        if (u1 < H || u1 >= W - H || u2 < H || u2 >= W - H) continue;
        const int c1 = o1 + u1, c2 = o2 + u2;
        if (c1 < 1 || c1 > L.N[a1] || c2 < 1 || c2 > L.N[a2]) continue;
        int g[3];
        g[d] = Ip; g[a1] = c1; g[a2] = c2;
        x.c[(g[0] + g[1] + g[2]) & 1][cidx(L, g[0], g[1], g[2])] = xs[e];
    }
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
// C,L,R,B,T,A,F order, linearSystemFunctions.H:42-49).  Optionally writes
// per-block partial sums of r^2 (deterministic order).
template <typename T, int OP, bool NORM>
__global__ __launch_bounds__(BX*BY) void k_residual(LevelDims L, OpView<T> op,
                                                    Pair<const T> x,
                                                    Pair<const T> b,
                                                    Pair<T> r, Region rg,
                                                    double* partials)
{
    __shared__ double sm[BX*BY/64];
    const int I = rg.S[0] + blockIdx.z;
    const int J = rg.S[1] + blockIdx.y*BY + threadIdx.y;
    const int p = (rg.S[2] >> 1) + blockIdx.x*BX + threadIdx.x;
    double acc = 0.0;
    if (J < rg.E[1])
    {
        const int sI = L.sI;
        const int row = (I*L.P[1] + J)*L.Hz;
        const int idx = row + p;
#pragma unroll
        for (int h = 0; h < 2; h++)
        {
            const int K = 2*p + h;
            if (K < rg.S[2] || K >= rg.E[2]) continue;
            const int colour = (I + J + K) & 1;
            const T* xc = x.c[colour];
            const T* xo = x.c[colour ^ 1];
            T a[7];
            coeffs<T, OP>(L, op, colour, idx, row, p, h, I, J, K, a);
            const T Ax =
                a[0]*xc[idx]
              + a[1]*xo[idx - sI]
              + a[2]*xo[idx + sI]
              + a[3]*xo[idx - L.Hz]
              + a[4]*xo[idx + L.Hz]
              + a[5]*xo[row + p - 1 + h]
              + a[6]*xo[row + p + h];
            const T res = b.c[colour][idx] - Ax;
            r.c[colour][idx] = res;
            if (NORM) acc += (double)res*(double)res;
        }
    }
    if (NORM)
    {
        const double s = block_sum(acc, sm);
        if (threadIdx.x == 0 && threadIdx.y == 0)
            partials[blockIdx.x + gridDim.x*(blockIdx.y + gridDim.y*blockIdx.z)]
                = s;
    }
}

// ---------------------------------------------------------------------------
// Scaled linear restriction (linearRestrictionScheme.C:103-160, weights
// :15-80): b_c = cvc * sum_q (w_q * r_q) * (1/cvf_q), q over the 8 children
// (z outer, x inner), w_q = ((fx*fy)*fz) zeroed when <= 1e-12.
template <typename T>
__global__ __launch_bounds__(BX*BY) void k_restrict(LevelDims Lf, LevelDims Lc,
                                                    TransferView<T> tv,
                                                    Pair<const T> rf,
                                                    Pair<T> bc)
{
    const int I = 1 + blockIdx.z;
    const int J = 1 + blockIdx.y*BY + threadIdx.y;
    const int p = blockIdx.x*BX + threadIdx.x;
    if (J > Lc.N[1]) return;
    const int R0 = Lc.R[0], R1 = Lc.R[1], R2 = Lc.R[2];
    const int ox = R0 > 1, oy = R1 > 1, oz = R2 > 1;
    const T ux = tv.ru[0][I], uy = tv.ru[1][J];
    const T fx[2] = {T(1.0) - ux, ux};
    const T fy[2] = {T(1.0) - uy, uy};
#pragma unroll
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K < 1 || K > Lc.N[2]) continue;
        const T uz = tv.ru[2][K];
        const T fz[2] = {T(1.0) - uz, uz};
        T v = T(0);
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
            if (w != T(0))
            {
                const int fi = (I - 1)*R0 + cx*ox + 1;
                const int fj = (J - 1)*R1 + cy*oy + 1;
                const int fk = (K - 1)*R2 + cz*oz + 1;
                T cvf = tv.fw[0][fi]*tv.fw[1][fj];
                cvf = cvf*tv.fw[2][fk];
                cvf = cvf > T(1e-16) ? cvf : T(1e-16);
                const T icvf = T(1.0)/cvf;
                const int fc = (fi + fj + fk) & 1;
                const T t = w*rf.c[fc][cidx(Lf, fi, fj, fk)];
                v = v + t*icvf;
            }
        }
        T cvc = tv.cw[0][I]*tv.cw[1][J];
        cvc = cvc*tv.cw[2][K];
        cvc = cvc > T(1e-16) ? cvc : T(1e-16);
        v = v*cvc;
        bc.c[(I + J + K) & 1][(I*Lc.P[1] + J)*Lc.Hz + p] = v;
    }
}

// Linear prolongation with plusEqOp (linearProlongationScheme.C:112-147,
// weights :12-110): x_f += sum_v w_v x_c(idx_v), v = (cx,cy,cz) with x
// fastest, idx = i/R + abc*off, off = i odd ? +1 : -1, w_v = (gx*gy)*gz with
// g = 1-u (abc 0) or u (abc 1); u = 1/2 on all axes when any u lies outside
// [0,1) (:89-93).
//
// One thread per pair of fine cells K = 2p, 2p+1.  Along z the two cells use
// the same two coarse indices (for R = 2 the odd cell's parent is the even
// cell's neighbour and vice versa; for R = 1 both are {2p, 2p+1}), so the 8
// coarse values are loaded once per pair and each cell sums them in its own
// parent/neighbour order.
template <typename T>
__global__ __launch_bounds__(BX*BY) void k_prolong(LevelDims Lf, LevelDims Lc,
                                                   TransferView<T> tv,
                                                   Pair<const T> xc,
                                                   Pair<T> xf)
{
    const int I = 1 + blockIdx.z;
    const int J = 1 + blockIdx.y*BY + threadIdx.y;
    const int p = blockIdx.x*BX + threadIdx.x;
    if (J > Lf.N[1]) return;
    const int K0 = 2*p, K1 = 2*p + 1;
    const bool v0 = K0 >= 1 && K0 <= Lf.N[2];
    const bool v1 = K1 >= 1 && K1 <= Lf.N[2];
    if (!v0 && !v1) return;
    const int i = I - 1, j = J - 1;
    const int offx = (i & 1) ? 1 : -1, offy = (j & 1) ? 1 : -1;
    const int ci[2] = {i/Lc.R[0] + 1, i/Lc.R[0] + 1 + offx};
    const int cj[2] = {j/Lc.R[1] + 1, j/Lc.R[1] + 1 + offy};
    // coarse z indices shared by both cells: zl < zh
    const int zl = Lc.R[2] == 1 ? 2*p : p;
    const int zh = zl + 1;
    T c[2][2][2];   // [cz: zl/zh][cy][cx]
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
    for (int bb = 0; bb < 2; bb++)
    {
        const int base = (ci[a]*Lc.P[1] + cj[bb])*Lc.Hz;
        const int par = ci[a] + cj[bb];
        c[0][bb][a] = xc.c[(par + zl) & 1][base + (zl >> 1)];
        c[1][bb][a] = xc.c[(par + zh) & 1][base + (zh >> 1)];
    }
    const T ux = tv.pu[0][I], uy = tv.pu[1][J];
    const int colour0 = (I + J + K0) & 1;
    const int idx = (I*Lf.P[1] + J)*Lf.Hz + p;
#pragma unroll
    for (int h = 0; h < 2; h++)
    {
        if (h == 0 ? !v0 : !v1) continue;
        const int K = 2*p + h;
        const int k = K - 1;
        // parent / neighbour along z as indices into c[0..1]
        const int offz = (k & 1) ? 1 : -1;
        const int bz = k/Lc.R[2] + 1;
        const int zpar = bz == zl ? 0 : 1;
        const int znb = (bz + offz) == zl ? 0 : 1;
        T u0 = ux, u1 = uy, u2 = tv.pu[2][K];
        const bool inside = !(u0 < T(0) || u0 >= T(1) || u1 < T(0) ||
                              u1 >= T(1) || u2 < T(0) || u2 >= T(1));
        if (!inside) { u0 = T(0.5); u1 = T(0.5); u2 = T(0.5); }
        const T gx[2] = {T(1.0) - u0, u0};
        const T gy[2] = {T(1.0) - u1, u1};
        const T gz[2] = {T(1.0) - u2, u2};
        const int zsel[2] = {zpar, znb};
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
            if (w != T(0))
            {
                const T t = w*c[zsel[cz]][cy][cx];
                value = value + t;
            }
        }
        T* dst = xf.c[colour0 ^ h];
        dst[idx] = dst[idx] + value;
    }
}

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
        val = x.c[(s[0] + s[1] + s[2]) & 1][cidx(L, s[0], s[1], s[2])];
    }
    else
    {
        for (int d = 0; d < 3; d++) s[d] = g[d] - bo[d];
        const T xi = x.c[(s[0] + s[1] + s[2]) & 1][cidx(L, s[0], s[1], s[2])];
        if (nout > 1)
            val = xi;
        else
        {
            int d = bo[0] != 0 ? 0 : (bo[1] != 0 ? 1 : 2);
            const int f = 2*d + (bo[d] > 0);
            const T H = T(level0 ? 1.0 : 0.0);
            if (L.bc[f] == NEU)
            {
                const T delta = op.dl[d][bo[d] > 0 ? L.N[d] + 1 : 1];
                val = xi + H*T(bv.v[f])/delta;
            }
            else
                val = H*T(2.0)*T(bv.v[f]) - xi;
        }
    }
    x.c[(I + J + K) & 1][cidx(L, I, J, K)] = val;
}

__device__ __forceinline__ bool shell_cell(const LevelDims& L, long t, int& I,
                                           int& J, int& K)
{
    const long P0 = L.P[0], P1 = L.P[1], P2 = L.P[2];
    const long nA = 2*P1*P2;
    const long nB = 2*(P0 - 2)*P2;
    const long nC = 2*(P0 - 2)*(P1 - 2);
    if (t < nA)
    {
        I = t < P1*P2 ? 0 : int(P0 - 1);
        const long u = t % (P1*P2);
        J = int(u/P2);
        K = int(u % P2);
        return true;
    }
    t -= nA;
    if (t < nB)
    {
        const long half = (P0 - 2)*P2;
        J = t < half ? 0 : int(P1 - 1);
        const long u = t % half;
        I = 1 + int(u/P2);
        K = int(u % P2);
        return true;
    }
    t -= nB;
    if (t < nC)
    {
        const long half = (P0 - 2)*(P1 - 2);
        K = t < half ? 0 : int(P2 - 1);
        const long u = t % half;
        I = 1 + int(u/(P1 - 2));
        J = 1 + int(u % (P1 - 2));
        return true;
    }
    return false;
}

__host__ __device__ inline long shell_count(const LevelDims& L)
{
    const long P0 = L.P[0], P1 = L.P[1], P2 = L.P[2];
    return 2*P1*P2 + 2*(P0 - 2)*P2 + 2*(P0 - 2)*(P1 - 2);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bc_fill(LevelDims L, OpView<T> op,
                                                 Pair<T> x, int mode,
                                                 int level0, BcVals bv)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    int I, J, K;
    if (!shell_cell(L, t, I, J, K)) return;
    ghost_cell(L, op, x, mode, level0, bv, I, J, K);
}

template <typename T>
__global__ void k_fill(T* a, T* b, long n, T v)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t < n) { a[t] = v; b[t] = v; }
}

// ---------------------------------------------------------------------------
// Layout conversions between the reference block layout and colour-split.
template <typename T>
__global__ __launch_bounds__(BX*BY) void k_ref_to_split(LevelDims L,
                                                        const double* ref,
                                                        Pair<T> out)
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
        out.c[(I + J + K) & 1][idx] =
            T(ref[((long)I*L.P[1] + J)*L.P[2] + K]);
    }
}

template <typename T>
__global__ __launch_bounds__(BX*BY) void k_split_to_ref(LevelDims L,
                                                        Pair<const T> in,
                                                        double* ref)
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
        ref[((long)I*L.P[1] + J)*L.P[2] + K] =
            double(in.c[(I + J + K) & 1][idx]);
    }
}

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
__global__ __launch_bounds__(BX*BY) void k_pin(LevelDims L, Pair<double> a7)
{
    const int I = 1 + blockIdx.z;
    const int J = 1 + blockIdx.y*BY + threadIdx.y;
    const int p = blockIdx.x*BX + threadIdx.x;
    if (J > L.N[1]) return;
    const int idx = (I*L.P[1] + J)*L.Hz + p;
    const int off[7][3] = {{0,0,0},{-1,0,0},{1,0,0},{0,-1,0},{0,1,0},
                           {0,0,-1},{0,0,1}};
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*p + h;
        if (K < 1 || K > L.N[2]) continue;
        double* a = a7.c[(I + J + K) & 1];
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
__global__ __launch_bounds__(BX*BY) void k_bc_source(LevelDims L,
                                                     OpView<double> op,
                                                     Pair<double> b, BcVals bv)
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
        double a[7];
        raw_coeffs<double, OP>(L, op, colour, idx, row, p, h, I, J, K, a);
        const bool on[6] = {I == 1, I == L.N[0], J == 1, J == L.N[1],
                            K == 1, K == L.N[2]};
        const int pc[3] = {I, J, K};
        double v = b.c[colour][idx];
        for (int f = 0; f < 6; f++)
        {
            if (!on[f] || L.bc[f] == PER) continue;
            const double g = a[f + 1];
            if (L.bc[f] == NEU)
            {
                const double delta = op.dl[f/2][pc[f/2] + (f & 1)];
                v = v - g*bv.v[f]/delta;
            }
            else
                v = v - g*2.0*bv.v[f];
        }
        b.c[colour][idx] = v;
    }
}

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

// ---------------------------------------------------------------------------
// Coarsest-level direct solve, one workgroup (APLU.C:91-128 with rhsSource,
// linearSystemAggregation.C:336-430).  The permutation is applied first and
// the unit-lower forward substitution runs column by column; for each row
// the subtractions still happen in column order, exactly as the row-wise
// LUBacksubstitute.  The upper back substitution is row-wise in one lane.
// The factors are staged in LDS when they fit.  Then the full ghost update
// of the level (APLU.C:126).
template <typename T, bool LDS>
__global__ __launch_bounds__(256) void k_coarse_solve(LevelDims L,
                                                      OpView<T> op, Pair<T> x,
                                                      Pair<const T> b,
                                                      const double* lug,
                                                      const int* piv, int n,
                                                      int level0, BcVals bv)
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
        double v = double(b.c[colour][idx]);
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
                const double t = a*double(x.c[colour ^ 1][cidx(L, g[0], g[1], g[2])]);
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
        x.c[(I + J + K) & 1][cidx(L, I, J, K)] = T(y[q]);
    }
    __syncthreads();
    const long ns = shell_count(L);
    for (long t = tid; t < ns; t += nt)
    {
        int I, J, K;
        if (shell_cell(L, t, I, J, K))
            ghost_cell(L, op, x, 1, level0, bv, I, J, K);
    }
}

// ---------------------------------------------------------------------------
// Reductions over interior cells (per-block partials, fixed order).
template <typename T, int MODE>
__global__ __launch_bounds__(BX*BY) void k_reduce_field(LevelDims L,
                                                        Pair<const T> f,
                                                        double* partials)
{
    __shared__ double sm[BX*BY/64];
    const int I = 1 + blockIdx.z;
    const int J = 1 + blockIdx.y*BY + threadIdx.y;
    const int p = blockIdx.x*BX + threadIdx.x;
    double acc = 0.0;
    if (J <= L.N[1])
    {
        const int idx = (I*L.P[1] + J)*L.Hz + p;
        for (int h = 0; h < 2; h++)
        {
            const int K = 2*p + h;
            if (K < 1 || K > L.N[2]) continue;
            const double v = double(f.c[(I + J + K) & 1][idx]);
            acc += MODE == 0 ? v*v : v;
        }
    }
    const double s = block_sum(acc, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partials[blockIdx.x + gridDim.x*(blockIdx.y + gridDim.y*blockIdx.z)] = s;
}

// Terms of normFactors (solver.C:35-46): (|r - b + Ay| + |b - Ay|)^2 with
// Ay = rowSum(A)*y, rowSum = C + L + R + B + T + A + F in that order.
template <typename T, int OP>
__global__ __launch_bounds__(BX*BY) void k_nf_terms(LevelDims L, OpView<T> op,
                                                    Pair<const T> r,
                                                    Pair<const T> b, double y,
                                                    double* partials)
{
    __shared__ double sm[BX*BY/64];
    const int I = 1 + blockIdx.z;
    const int J = 1 + blockIdx.y*BY + threadIdx.y;
    const int p = blockIdx.x*BX + threadIdx.x;
    double acc = 0.0;
    if (J <= L.N[1])
    {
        const int row = (I*L.P[1] + J)*L.Hz;
        const int idx = row + p;
        for (int h = 0; h < 2; h++)
        {
            const int K = 2*p + h;
            if (K < 1 || K > L.N[2]) continue;
            const int colour = (I + J + K) & 1;
            T a[7];
            coeffs<T, OP>(L, op, colour, idx, row, p, h, I, J, K, a);
            double rs = double(a[0]);
            for (int s = 1; s < 7; s++) rs = rs + double(a[s]);
            const double Ay = rs*y;
            const double rv = double(r.c[colour][idx]);
            const double bv = double(b.c[colour][idx]);
            const double t = fabs((rv - bv) + Ay) + fabs(bv - Ay);
            acc += t*t;
        }
    }
    const double s = block_sum(acc, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partials[blockIdx.x + gridDim.x*(blockIdx.y + gridDim.y*blockIdx.z)] = s;
}

__global__ __launch_bounds__(1024) void k_reduce_final(const double* partials,
                                                       int n, double* out)
{
    __shared__ double sm[16];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int i = threadIdx.x;
    for (; i + 3*1024 < n; i += 4*1024)
    {
        a0 += partials[i];
        a1 += partials[i + 1024];
        a2 += partials[i + 2*1024];
        a3 += partials[i + 3*1024];
    }
    for (; i < n; i += 1024) a0 += partials[i];
    const double s = block_sum((a0 + a1) + (a2 + a3), sm);
    if (threadIdx.x == 0) out[0] = s;
}

// q = A p (row product in the C,L,R,B,T,A,F order), interior cells.
template <int OP>
__global__ __launch_bounds__(BX*BY) void k_apply(LevelDims L,
                                                 OpView<double> op,
                                                 Pair<const double> p,
                                                 Pair<double> q)
{
    const int I = 1 + blockIdx.z;
    const int J = 1 + blockIdx.y*BY + threadIdx.y;
    const int pp = blockIdx.x*BX + threadIdx.x;
    if (J > L.N[1]) return;
    const int row = (I*L.P[1] + J)*L.Hz;
    const int idx = row + pp;
    for (int h = 0; h < 2; h++)
    {
        const int K = 2*pp + h;
        if (K < 1 || K > L.N[2]) continue;
        const int colour = (I + J + K) & 1;
        const double* xc = p.c[colour];
        const double* xo = p.c[colour ^ 1];
        double a[7];
        coeffs<double, OP>(L, op, colour, idx, row, pp, h, I, J, K, a);
        q.c[colour][idx] =
            a[0]*xc[idx]
          + a[1]*xo[idx - L.sI]
          + a[2]*xo[idx + L.sI]
          + a[3]*xo[idx - L.Hz]
          + a[4]*xo[idx + L.Hz]
          + a[5]*xo[row + pp - 1 + h]
          + a[6]*xo[row + pp + h];
    }
}

// Up to three dot products over interior cells in one pass; per-block
// partials at partials[d*nblocks + block].
__global__ __launch_bounds__(BX*BY) void k_dots(LevelDims L, int nd,
                                                Pair<const double> a0,
                                                Pair<const double> b0,
                                                Pair<const double> a1,
                                                Pair<const double> b1,
                                                Pair<const double> a2,
                                                Pair<const double> b2,
                                                double* partials)
{
    __shared__ double sm[BX*BY/64];
    const int I = 1 + blockIdx.z;
    const int J = 1 + blockIdx.y*BY + threadIdx.y;
    const int pp = blockIdx.x*BX + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (J <= L.N[1])
    {
        const int idx = (I*L.P[1] + J)*L.Hz + pp;
        for (int h = 0; h < 2; h++)
        {
            const int K = 2*pp + h;
            if (K < 1 || K > L.N[2]) continue;
            const int c = (I + J + K) & 1;
            acc[0] += a0.c[c][idx]*b0.c[c][idx];
            if (nd > 1) acc[1] += a1.c[c][idx]*b1.c[c][idx];
            if (nd > 2) acc[2] += a2.c[c][idx]*b2.c[c][idx];
        }
    }
    const int nb = gridDim.x*gridDim.y*gridDim.z;
    const int blk = blockIdx.x + gridDim.x*(blockIdx.y + gridDim.y*blockIdx.z);
    for (int d = 0; d < nd; d++)
    {
        const double s = block_sum(acc[d], sm);
        if (threadIdx.x == 0 && threadIdx.y == 0) partials[d*nb + blk] = s;
        __syncthreads();
    }
}

// x += alpha p; r -= alpha q (interior) and per-block partial sums of r^2.
__global__ __launch_bounds__(BX*BY) void k_cg_update(LevelDims L, double alpha,
                                                     Pair<double> x,
                                                     Pair<double> r,
                                                     Pair<const double> p,
                                                     Pair<const double> q,
                                                     double* partials)
{
    __shared__ double sm[BX*BY/64];
    const int I = 1 + blockIdx.z;
    const int J = 1 + blockIdx.y*BY + threadIdx.y;
    const int pp = blockIdx.x*BX + threadIdx.x;
    double acc = 0.0;
    if (J <= L.N[1])
    {
        const int idx = (I*L.P[1] + J)*L.Hz + pp;
        for (int h = 0; h < 2; h++)
        {
            const int K = 2*pp + h;
            if (K < 1 || K > L.N[2]) continue;
            const int c = (I + J + K) & 1;
            x.c[c][idx] = x.c[c][idx] + alpha*p.c[c][idx];
            const double rn = r.c[c][idx] - alpha*q.c[c][idx];
            r.c[c][idx] = rn;
            acc += rn*rn;
        }
    }
    const double s = block_sum(acc, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partials[blockIdx.x + gridDim.x*(blockIdx.y + gridDim.y*blockIdx.z)] = s;
}

// p = z + beta p over the whole padded arrays (ghosts refreshed later).
__global__ void k_xpby(long n, Pair<const double> z, double beta,
                       Pair<double> p)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t < n)
    {
        p.c[0][t] = z.c[0][t] + beta*p.c[0][t];
        p.c[1][t] = z.c[1][t] + beta*p.c[1][t];
    }
}

template <typename T>
__global__ void k_copy(long n, Pair<const T> a, Pair<T> b)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t < n) { b.c[0][t] = a.c[0][t]; b.c[1][t] = a.c[1][t]; }
}

__global__ void k_to_double(long n, Pair<const float> a, Pair<double> b)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t < n) { b.c[0][t] = double(a.c[0][t]); b.c[1][t] = double(a.c[1][t]); }
}

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

__global__ __launch_bounds__(BX*BY) void k_add_correction(LevelDims L,
                                                          Pair<double> x,
                                                          Pair<const float> e)
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
        x.c[c][idx] = x.c[c][idx] + double(e.c[c][idx]);
    }
}

__global__ void k_to_float(long n, Pair<const double> in, Pair<float> out)
{
    const long t = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t < n)
    {
        out.c[0][t] = float(in.c[0][t]);
        out.c[1][t] = float(in.c[1][t]);
    }
}

template <typename T>
__global__ void k_set_cell(Pair<T> f, long idx, int colour, T v)
{
    f.c[colour][idx] = v;
}

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

BcVals bcvals(const double* v)
{
    BcVals b;
    for (int f = 0; f < 6; f++) b.v[f] = v ? v[f] : 0.0;
    return b;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
template <typename T>
void launch_rbgs(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                 Pair<const T> b, int colour, const Region& rg, hipStream_t s)
{
    if (rg.E[0] <= rg.S[0] || rg.E[1] <= rg.S[1] || rg.E[2] <= rg.S[2]) return;
    const int nK = rg.E[2] - rg.S[2];
    const dim3 g(((nK + 1)/2 + 1 + BX - 1)/BX, (rg.E[1] - rg.S[1] + BY - 1)/BY,
                 rg.E[0] - rg.S[0]);
    T* xc = x.c[colour];
    const T* xo = x.c[colour ^ 1];
    const T* bc = b.c[colour];
    switch (op.kind)
    {
    case OP_STENCIL:
        hipLaunchKernelGGL((k_rbgs<T, OP_STENCIL>), g, BLK, 0, s, L, op, xc,
                           xo, bc, colour, rg);
        break;
    case OP_CONST:
        hipLaunchKernelGGL((k_rbgs<T, OP_CONST>), g, BLK, 0, s, L, op, xc, xo,
                           bc, colour, rg);
        break;
    default:
        hipLaunchKernelGGL((k_rbgs<T, OP_FACES>), g, BLK, 0, s, L, op, xc, xo,
                           bc, colour, rg);
        break;
    }
}

template <typename T>
void launch_rbgs_slabs(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                       Pair<const T> b, int d, const int* planes, int nslab,
                       const int* colours, int nphase, hipStream_t s)
{
    SlabArgs sa;
    sa.d = d;
    sa.plane[0] = planes[0];
    sa.plane[1] = nslab > 1 ? planes[1] : planes[0];
    sa.nphase = nphase;
    for (int z = 0; z < 2; z++)
        for (int i = 0; i < 4; i++)
            sa.colours[z][i] = i < nphase ? colours[(nslab > 1 ? z : 0)*4 + i] : 0;
    const int a1 = d == 0 ? 1 : 0;
    const int a2 = d == 2 ? 1 : 2;
    const int n1 = L.N[a1], n2 = L.N[a2];
    sa.tt = (n1*n2 >= 128*128) ? 16 : 8;
    const int W = sa.tt + 2*nphase;
    const dim3 g((n2 + sa.tt - 1)/sa.tt, (n1 + sa.tt - 1)/sa.tt, nslab);
    const size_t lds = size_t(W)*W*sizeof(T);
    switch (op.kind)
    {
    case OP_STENCIL:
        hipLaunchKernelGGL((k_rbgs_slabs<T, OP_STENCIL>), g, dim3(256), lds, s,
                           L, op, x, b, sa);
        break;
    case OP_CONST:
        hipLaunchKernelGGL((k_rbgs_slabs<T, OP_CONST>), g, dim3(256), lds, s,
                           L, op, x, b, sa);
        break;
    default:
        hipLaunchKernelGGL((k_rbgs_slabs<T, OP_FACES>), g, dim3(256), lds, s,
                           L, op, x, b, sa);
        break;
    }
}

template <typename T>
void launch_jacobi(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                   Pair<const T> b, Pair<T> tmp, const Region& rg,
                   hipStream_t s)
{
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
void launch_residual(const LevelDims& L, const OpView<T>& op, Pair<const T> x,
                     Pair<const T> b, Pair<T> r, const Region& rg,
                     double* partials, int* nblocks, hipStream_t s)
{
    const dim3 g = grid3(rg);
    if (nblocks) *nblocks = int(g.x*g.y*g.z);
    if (rg.E[0] <= rg.S[0] || rg.E[1] <= rg.S[1] || rg.E[2] <= rg.S[2]) return;
#define BMG_RES(OPK, NRM)                                                     \
    hipLaunchKernelGGL((k_residual<T, OPK, NRM>), g, BLK, 0, s, L, op, x, b, \
                       r, rg, partials)
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
    hipLaunchKernelGGL((k_restrict<T>), grid_interior(Lc), BLK, 0, s, Lf, Lc,
                       tv, rf, bc);
}

template <typename T>
void launch_prolong(const LevelDims& Lf, const LevelDims& Lc,
                    const TransferView<T>& tv, Pair<const T> xc, Pair<T> xf,
                    hipStream_t s)
{
    hipLaunchKernelGGL((k_prolong<T>), grid_interior(Lf), BLK, 0, s, Lf, Lc,
                       tv, xc, xf);
}

template <typename T>
void launch_bc_fill(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                    int mode, int level0, const double* bcval, hipStream_t s)
{
    const long n = shell_count(L);
    hipLaunchKernelGGL((k_bc_fill<T>), dim3((n + 255)/256), dim3(256), 0, s, L,
                       op, x, mode, level0, bcvals(bcval));
}

template <typename T>
void launch_fill(Pair<T> x, long ncol, T v, hipStream_t s)
{
    hipLaunchKernelGGL((k_fill<T>), dim3((ncol + 255)/256), dim3(256), 0, s,
                       x.c[0], x.c[1], ncol, v);
}

template <typename T>
void launch_ref_to_split(const LevelDims& L, const double* ref, Pair<T> out,
                         hipStream_t s)
{
    hipLaunchKernelGGL((k_ref_to_split<T>), grid_padded(L), BLK, 0, s, L, ref,
                       out);
}

template <typename T>
void launch_split_to_ref(const LevelDims& L, Pair<const T> in, double* ref,
                         hipStream_t s)
{
    hipLaunchKernelGGL((k_split_to_ref<T>), grid_padded(L), BLK, 0, s, L, in,
                       ref);
}

void launch_stencil_ref_to_split(const LevelDims& L, const double* ref7,
                                 Pair<double> a7, hipStream_t s)
{
    hipLaunchKernelGGL(k_stencil_ref_to_split, grid_padded(L), BLK, 0, s, L,
                       ref7, a7);
}

void launch_stencil_split_to_ref(const LevelDims& L, Pair<const double> a7,
                                 double* ref7, hipStream_t s)
{
    hipLaunchKernelGGL(k_stencil_split_to_ref, grid_padded(L), BLK, 0, s, L,
                       a7, ref7);
}

template <typename T>
void launch_materialize_stencil(const LevelDims& L, const OpView<T>& op,
                                Pair<T> a7, hipStream_t s)
{
    if (op.kind == OP_CONST)
        hipLaunchKernelGGL((k_materialize<T, OP_CONST>), grid_interior(L), BLK,
                           0, s, L, op, a7);
    else if (op.kind == OP_FACES)
        hipLaunchKernelGGL((k_materialize<T, OP_FACES>), grid_interior(L), BLK,
                           0, s, L, op, a7);
}

void launch_pin(const LevelDims& L, Pair<double> a7, hipStream_t s)
{
    hipLaunchKernelGGL(k_pin, grid_interior(L), BLK, 0, s, L, a7);
}

void launch_eliminate(const LevelDims& L, Pair<double> a7, hipStream_t s)
{
    hipLaunchKernelGGL(k_eliminate, grid_interior(L), BLK, 0, s, L, a7);
}

void launch_bc_source(const LevelDims& L, const OpView<double>& op,
                      Pair<double> b, const double* bcval, hipStream_t s)
{
    if (op.kind == OP_STENCIL)
        hipLaunchKernelGGL((k_bc_source<OP_STENCIL>), grid_interior(L), BLK, 0,
                           s, L, op, b, bcvals(bcval));
    else if (op.kind == OP_CONST)
        hipLaunchKernelGGL((k_bc_source<OP_CONST>), grid_interior(L), BLK, 0,
                           s, L, op, b, bcvals(bcval));
    else
        hipLaunchKernelGGL((k_bc_source<OP_FACES>), grid_interior(L), BLK, 0,
                           s, L, op, b, bcvals(bcval));
}

void launch_lambda_restrict(const LevelDims& Lf, const LevelDims& Lc,
                            const double* fw0, const double* fw1,
                            const double* fw2, const double* lamf,
                            double* lamc, int fd, hipStream_t s)
{
    const long n = long(Lc.N[0] + (fd == 0))*(Lc.N[1] + (fd == 1))*
                   (Lc.N[2] + (fd == 2));
    hipLaunchKernelGGL(k_lambda_restrict, dim3((n + 255)/256), dim3(256), 0, s,
                       Lf, Lc, fw0, fw1, fw2, lamf, lamc, fd);
}

void launch_face_coeffs(const LevelDims& L, const double* w0, const double* w1,
                        const double* w2, const double* dl, const double* lam,
                        int fd, Pair<double> wf, hipStream_t s)
{
    const long n = long(L.N[0] + (fd == 0))*(L.N[1] + (fd == 1))*
                   (L.N[2] + (fd == 2));
    hipLaunchKernelGGL(k_face_coeffs, dim3((n + 255)/256), dim3(256), 0, s, L,
                       w0, w1, w2, dl, lam, fd, wf);
}

template <typename T>
void launch_coarse_solve(const LevelDims& L, const OpView<T>& op, Pair<T> x,
                         Pair<const T> b, const double* lu, const int* piv,
                         int n, int level0, const double* bcval,
                         hipStream_t s)
{
    if (size_t(n)*(n + 1)*sizeof(double) <= 56*1024)
        hipLaunchKernelGGL((k_coarse_solve<T, true>), dim3(1), dim3(256),
                           size_t(n)*(n + 1)*sizeof(double), s, L, op, x, b,
                           lu, piv, n, level0, bcvals(bcval));
    else
        hipLaunchKernelGGL((k_coarse_solve<T, false>), dim3(1), dim3(256),
                           size_t(n)*sizeof(double), s, L, op, x, b, lu, piv,
                           n, level0, bcvals(bcval));
}

template <typename T>
void launch_sum_sq(const LevelDims& L, Pair<const T> f, double* partials,
                   int* nblocks, hipStream_t s)
{
    const dim3 g = grid_interior(L);
    *nblocks = int(g.x*g.y*g.z);
    hipLaunchKernelGGL((k_reduce_field<T, 0>), g, BLK, 0, s, L, f, partials);
}

template <typename T>
void launch_sum(const LevelDims& L, Pair<const T> f, double* partials,
                int* nblocks, hipStream_t s)
{
    const dim3 g = grid_interior(L);
    *nblocks = int(g.x*g.y*g.z);
    hipLaunchKernelGGL((k_reduce_field<T, 1>), g, BLK, 0, s, L, f, partials);
}

template <typename T>
void launch_norm_factor_terms(const LevelDims& L, const OpView<T>& op,
                              Pair<const T> r, Pair<const T> b, double y,
                              double* partials, int* nblocks, hipStream_t s)
{
    const dim3 g = grid_interior(L);
    *nblocks = int(g.x*g.y*g.z);
    if (op.kind == OP_STENCIL)
        hipLaunchKernelGGL((k_nf_terms<T, OP_STENCIL>), g, BLK, 0, s, L, op, r,
                           b, y, partials);
    else if (op.kind == OP_CONST)
        hipLaunchKernelGGL((k_nf_terms<T, OP_CONST>), g, BLK, 0, s, L, op, r, b,
                           y, partials);
    else
        hipLaunchKernelGGL((k_nf_terms<T, OP_FACES>), g, BLK, 0, s, L, op, r, b,
                           y, partials);
}

void launch_reduce_final(const double* partials, int n, double* out,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, s, partials, n,
                       out);
}

void launch_diag(const LevelDims& L, Pair<const double> a7, Pair<double> x,
                 Pair<const double> b, hipStream_t s)
{
    hipLaunchKernelGGL(k_diag, grid_interior(L), BLK, 0, s, L, a7, x, b);
}

void launch_apply(const LevelDims& L, const OpView<double>& op,
                  Pair<const double> p, Pair<double> q, hipStream_t s)
{
    if (op.kind == OP_STENCIL)
        hipLaunchKernelGGL((k_apply<OP_STENCIL>), grid_interior(L), BLK, 0, s,
                           L, op, p, q);
    else if (op.kind == OP_CONST)
        hipLaunchKernelGGL((k_apply<OP_CONST>), grid_interior(L), BLK, 0, s, L,
                           op, p, q);
    else
        hipLaunchKernelGGL((k_apply<OP_FACES>), grid_interior(L), BLK, 0, s, L,
                           op, p, q);
}

int launch_dots(const LevelDims& L, int nd, const Pair<const double>* ab,
                double* partials, hipStream_t s)
{
    const dim3 g = grid_interior(L);
    Pair<const double> z{{nullptr, nullptr}};
    hipLaunchKernelGGL(k_dots, g, BLK, 0, s, L, nd, ab[0], ab[1],
                       nd > 1 ? ab[2] : z, nd > 1 ? ab[3] : z,
                       nd > 2 ? ab[4] : z, nd > 2 ? ab[5] : z, partials);
    return int(g.x*g.y*g.z);
}

int launch_cg_update(const LevelDims& L, double alpha, Pair<double> x,
                     Pair<double> r, Pair<const double> p,
                     Pair<const double> q, double* partials, hipStream_t s)
{
    const dim3 g = grid_interior(L);
    hipLaunchKernelGGL(k_cg_update, g, BLK, 0, s, L, alpha, x, r, p, q,
                       partials);
    return int(g.x*g.y*g.z);
}

void launch_xpby(long n, Pair<const double> z, double beta, Pair<double> p,
                 hipStream_t s)
{
    hipLaunchKernelGGL(k_xpby, dim3((n + 255)/256), dim3(256), 0, s, n, z,
                       beta, p);
}

template <typename T>
void launch_copy(long n, Pair<const T> a, Pair<T> b, hipStream_t s)
{
    hipLaunchKernelGGL((k_copy<T>), dim3((n + 255)/256), dim3(256), 0, s, n,
                       a, b);
}
template void launch_copy<double>(long, Pair<const double>, Pair<double>,
                                  hipStream_t);

void launch_to_double(long n, Pair<const float> a, Pair<double> b,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_to_double, dim3((n + 255)/256), dim3(256), 0, s, n,
                       a, b);
}

void launch_add_correction(const LevelDims& L, Pair<double> x,
                           Pair<const float> e, hipStream_t s)
{
    hipLaunchKernelGGL(k_add_correction, grid_interior(L), BLK, 0, s, L, x, e);
}

void launch_to_float(long ncol, Pair<const double> in, Pair<float> out,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_to_float, dim3((ncol + 255)/256), dim3(256), 0, s,
                       ncol, in, out);
}

template <typename T>
void launch_set_cell(Pair<T> f, long idx, int colour, T v, hipStream_t s)
{
    hipLaunchKernelGGL((k_set_cell<T>), dim3(1), dim3(1), 0, s, f, idx, colour,
                       v);
}

#define BMG_INST(T)                                                           \
    template void launch_rbgs<T>(const LevelDims&, const OpView<T>&, Pair<T>, \
                                 Pair<const T>, int, const Region&,           \
                                 hipStream_t);                                \
    template void launch_rbgs_slabs<T>(const LevelDims&, const OpView<T>&,    \
                                       Pair<T>, Pair<const T>, int,           \
                                       const int*, int, const int*, int,      \
                                       hipStream_t);                          \
    template void launch_jacobi<T>(const LevelDims&, const OpView<T>&,        \
                                   Pair<T>, Pair<const T>, Pair<T>,           \
                                   const Region&, hipStream_t);               \
    template void launch_residual<T>(const LevelDims&, const OpView<T>&,      \
                                     Pair<const T>, Pair<const T>, Pair<T>,   \
                                     const Region&, double*, int*,            \
                                     hipStream_t);                            \
    template void launch_restrict<T>(const LevelDims&, const LevelDims&,      \
                                     const TransferView<T>&, Pair<const T>,   \
                                     Pair<T>, hipStream_t);                   \
    template void launch_prolong<T>(const LevelDims&, const LevelDims&,       \
                                    const TransferView<T>&, Pair<const T>,    \
                                    Pair<T>, hipStream_t);                    \
    template void launch_bc_fill<T>(const LevelDims&, const OpView<T>&,       \
                                    Pair<T>, int, int, const double*,         \
                                    hipStream_t);                             \
    template void launch_fill<T>(Pair<T>, long, T, hipStream_t);              \
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
                                         const double*, hipStream_t);         \
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
