// HIP kernels of the learned preconditioner: a 3-D convolutional U-Net in
// BF16 with FP32 accumulation on the gfx950 matrix cores.
//
// Activations are channels-last ("NDHWC" with N = 1): element (m, c) of a
// tensor with row stride ld sits at m*ld + c, m = (z*H + y)*W + x.  A
// channel slice of a wider tensor is addressed by (base + offset, ld), which
// is how the U-Net concatenates skip connections without copies.
//
// Convolutions are implicit GEMMs over M = voxels, N = output channels and
// K = 27*Cin (tap-major: k = t*Cin + c, t = (dz+1)*9 + (dy+1)*3 + (dx+1)),
// zero padding at the domain boundary.  Weights are stored [Cout][Kp] with
// Kp = K rounded up to 32 and zero-filled.  The matrix instruction is
// v_mfma_f32_16x16x32_bf16: per wavefront a 16 x 16 FP32 tile of D += A B,
// A 16 x 32 (lane l: row l%16, k = 8*(l/16)..+7), B 32 x 16 (lane l: column
// l%16, k = 8*(l/16)..+7), D (lane l: column l%16, rows 4*(l/16)..+3).
#include "unet_internal.h"

namespace bmg {
namespace unet {

namespace {

typedef __attribute__((__vector_size__(8*sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4*sizeof(float)))) float floatx4;

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = __bf16(0.0f);
    return v;
}

__device__ __forceinline__ float bf2f(bf16 v) { return float(v); }

// Voxel m of a D x H x W grid shifted by tap t; -1 outside the grid.
__device__ __forceinline__ int tap_voxel(const Grid3& g, int z, int y, int x,
                                         int t)
{
    const int dz = t/9 - 1, dy = (t/3) % 3 - 1, dx = t % 3 - 1;
    const int zz = z + dz, yy = y + dy, xx = x + dx;
    if (zz < 0 || zz >= g.D || yy < 0 || yy >= g.H || xx < 0 || xx >= g.W)
        return -1;
    return (zz*g.H + yy)*g.W + xx;
}

// ---------------------------------------------------------------------------
// Forward convolution, 3x3x3 (or 1x1x1 when taps == 1), optional bias and
// ReLU.  A workgroup of 4 wavefronts computes 64 voxels x all output
// channels (NT = Cout/16 tiles per wavefront, Cout <= 192); each wavefront
// owns 16 voxels.  The A fragment (8 consecutive k of one voxel, i.e. 8
// channels of one tap, Cin % 8 == 0) is loaded once per K step and reused
// for every output tile.
template <int NT>
__global__ __launch_bounds__(256) void k_conv_fwd(ConvArgs a)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int M = a.g.D*a.g.H*a.g.W;
    const int m0 = (blockIdx.x*4 + wave)*16;
    if (m0 >= M) return;
    const int row = lane & 15;
    const int kq = (lane >> 4)*8;
    const int m = m0 + row;
    const bool mok = m < M;
    const int x = mok ? m % a.g.W : 0;
    const int y = mok ? (m/a.g.W) % a.g.H : 0;
    const int z = mok ? m/(a.g.W*a.g.H) : 0;
    const int K = a.taps*a.cin;
    floatx4 acc[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < a.kp; k0 += 32)
    {
        const int k = k0 + kq;
        bf16x8 av = zero8();
        if (mok && k < K)
        {
            const int t = k/a.cin, c = k - t*a.cin;
            const int v = a.taps == 1 ? m : tap_voxel(a.g, z, y, x, t);
            if (v >= 0)
                av = *reinterpret_cast<const bf16x8*>(a.x + size_t(v)*a.ldx + c);
        }
#pragma unroll
        for (int j = 0; j < NT; j++)
        {
            const int n = j*16 + row;
            const bf16x8 bv = *reinterpret_cast<const bf16x8*>(
                a.w + size_t(n)*a.kp + k);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[j], 0,
                                                            0, 0);
        }
    }
    // D: lane holds column n = j*16 + row, voxels m0 + 4*(lane/16) + r
#pragma unroll
    for (int j = 0; j < NT; j++)
    {
        const int n = j*16 + row;
        if (n >= a.cout) continue;
        const float bias = a.bias ? a.bias[n] : 0.0f;
#pragma unroll
        for (int r = 0; r < 4; r++)
        {
            const int mm = m0 + 4*(lane >> 4) + r;
            if (mm >= M) continue;
            float v = acc[j][r] + bias;
            if (a.relu && v < 0.0f) v = 0.0f;
            a.y[size_t(mm)*a.ldy + n] = bf16(v);
        }
    }
}

// LDS-staged forward convolution, 3x3x3.  A workgroup of 4 wavefronts
// computes an output tile of 16 x TY x TZ voxels (x fastest): tile row
// r = 0 .. 4*MT-1 is (y, z) = (r % TY, r / TY) with 16 consecutive x, and
// wavefront w owns rows w*MT .. w*MT+MT-1, one 16-voxel M tile each, times
// all NT output-channel tiles.  The input tile with its one-voxel halo,
// 18 x (TY+2) x (TZ+2) voxels of CH channels, is staged in LDS once per
// channel chunk (CH = Cin when Cin <= 48, else 32-channel chunks, Cin % 32
// == 0), zero outside the domain; every tap's A fragment is then a 16-byte
// LDS read, shared by the NT MFMAs of the M tile, and each B fragment (the
// weights of one k step) is read once per wavefront and shared by its MT
// tiles.  K order within a chunk of all channels: k = t*Cin + c as the
// weights are stored; with chunks, tap-major within the chunk (k = t*Cin +
// chunk*32 + 0..31), the same products summed in another order.
template <int MT>
struct ConvTile {
    static constexpr int TY = MT == 4 ? 8 : (MT == 2 ? 4 : 2);
    static constexpr int TZ = 4*MT/TY;
    static constexpr int HX = 18, HY = TY + 2, HZ = TZ + 2;
    static constexpr int HV = HX*HY*HZ;
};

// CIN > 0: all CIN channels staged at once, the k loop unrolled with each
// step's tap and channel known at compile time per lane group; CIN = 0:
// 32-channel chunks of a runtime channel count (a multiple of 32), the 27
// taps unrolled.  Consecutive k steps alternate between two accumulator
// sets (AS = 2 when the tiles leave registers for it) so that dependent
// MFMAs do not follow each other.
template <int NT, int MT, int CIN>
__global__ __launch_bounds__(256) void k_conv_fwd_lds(ConvArgs a)
{
    using Tl = ConvTile<MT>;
    constexpr int AS = NT*MT <= 8 ? 2 : 1;
    constexpr int CH = CIN > 0 ? CIN : 32;
    // LDS voxel stride in elements (padded by 8: 16-byte reads of 16
    // consecutive voxels spread over the banks), known at compile time so
    // that every tap offset is a constant
    constexpr int vs = CH + 8;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    bf16* lds = reinterpret_cast<bf16*>(lds_raw);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int row = lane & 15;
    const int g = lane >> 4;
    const int kq = g*8;
    const int W = a.g.W, H = a.g.H, D = a.g.D;
    const int ntx = (W + 15)/16, nty = (H + Tl::TY - 1)/Tl::TY;
    const int b = blockIdx.x;
    const int x0 = (b % ntx)*16;
    const int y0 = ((b/ntx) % nty)*Tl::TY;
    const int z0 = (b/(ntx*nty))*Tl::TZ;
    const int nchunk = a.cin/CH;
    constexpr int parts = CH/8;
    floatx4 acc[AS][MT][NT];
#pragma unroll
    for (int h = 0; h < AS; h++)
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
        for (int j = 0; j < NT; j++)
            acc[h][i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    // LDS element offset of this lane's voxel in tile row wave*MT + i (tap
    // offsets 0, the corner of its 3x3x3 stencil in the halo tile)
    int base[MT];
#pragma unroll
    for (int i = 0; i < MT; i++)
    {
        const int r = wave*MT + i;
        base[i] = (((r/Tl::TY)*Tl::HY + r % Tl::TY)*Tl::HX + row)*vs;
    }
    const bf16* wrow[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) wrow[j] = a.w + size_t(j*16 + row)*a.kp + kq;
    // halo-tile offset of tap t, in voxels
    auto tap_off = [](int t)
    {
        return (t/9)*Tl::HY*Tl::HX + ((t/3) % 3)*Tl::HX + t % 3;
    };
    for (int c0 = 0; c0 < nchunk; c0++)
    {
        if (c0 > 0) __syncthreads();
        for (int u = threadIdx.x; u < Tl::HV*parts; u += 256)
        {
            const int v = u/parts, part = u - v*parts;
            const int hx = v % Tl::HX, hy = (v/Tl::HX) % Tl::HY;
            const int hz = v/(Tl::HX*Tl::HY);
            const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
            bf16x8 val = zero8();
            if (gx >= 0 && gx < W && gy >= 0 && gy < H && gz >= 0 && gz < D)
                val = *reinterpret_cast<const bf16x8*>(
                    a.x + size_t((gz*H + gy)*W + gx)*a.ldx + c0*CH + part*8);
            *reinterpret_cast<bf16x8*>(lds + v*vs + part*8) = val;
        }
        __syncthreads();
        if constexpr (CIN > 0)
        {
            constexpr int K = 27*CIN;
            constexpr int NS = (K + 31)/32;
#pragma unroll
            for (int st = 0; st < NS; st++)
            {
                // this lane group's k = 32 st + 8 g: tap and channel
                int off = -1;
#pragma unroll
                for (int gg = 0; gg < 4; gg++)
                {
                    const int k = st*32 + gg*8;
                    if (k < K && g == gg)
                        off = tap_off(k/CIN)*vs + (k % CIN);
                }
                bf16x8 bv[NT];
#pragma unroll
                for (int j = 0; j < NT; j++)
                    bv[j] = *reinterpret_cast<const bf16x8*>(wrow[j] + st*32);
#pragma unroll
                for (int i = 0; i < MT; i++)
                {
                    bf16x8 av = zero8();
                    if (off >= 0)
                        av = *reinterpret_cast<const bf16x8*>(
                            lds + base[i] + off);
#pragma unroll
                    for (int j = 0; j < NT; j++)
                        acc[st % AS][i][j] =
                            __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                av, bv[j], acc[st % AS][i][j], 0, 0, 0);
                }
            }
        }
        else
        {
#pragma unroll
            for (int t = 0; t < 27; t++)
            {
                const int off = tap_off(t)*vs + kq;
                const int k = t*a.cin + c0*CH;
                bf16x8 bv[NT];
#pragma unroll
                for (int j = 0; j < NT; j++)
                    bv[j] = *reinterpret_cast<const bf16x8*>(wrow[j] + k);
#pragma unroll
                for (int i = 0; i < MT; i++)
                {
                    const bf16x8 av = *reinterpret_cast<const bf16x8*>(
                        lds + base[i] + off);
#pragma unroll
                    for (int j = 0; j < NT; j++)
                        acc[t % AS][i][j] =
                            __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                av, bv[j], acc[t % AS][i][j], 0, 0, 0);
                }
            }
        }
    }
    // D of tile (i, j): lane holds channel n = j*16 + row for the voxels
    // x = x0 + 4*(lane/16) + q, q = 0..3, of tile row wave*MT + i.  Each
    // 16 x 16 tile goes through the wavefront's own 512 bytes of LDS (the
    // input tile is no longer read) so that a lane stores 8 channels of one
    // voxel, 16 bytes; every wavefront passes the same barriers.
    __syncthreads();
    bf16* tb = lds + wave*256;
#pragma unroll
    for (int i = 0; i < MT; i++)
    {
        const int r = wave*MT + i;
        const int yy = y0 + r % Tl::TY, zz = z0 + r/Tl::TY;
#pragma unroll
        for (int j = 0; j < NT; j++)
        {
            const int n = j*16 + row;
            const float bias = a.bias ? a.bias[n] : 0.0f;
#pragma unroll
            for (int q = 0; q < 4; q++)
            {
                float v = acc[0][i][j][q] + bias;
                if (AS == 2) v = acc[0][i][j][q] + acc[AS - 1][i][j][q] + bias;
                if (a.relu && v < 0.0f) v = 0.0f;
                tb[(4*g + q)*16 + row] = bf16(v);
            }
            __syncthreads();
            // lanes 0..31: voxel x0 + lane/2, channels j*16 + 8*(lane&1)
            if (lane < 32 && yy < H && zz < D)
            {
                const int xx = x0 + (lane >> 1);
                const int n0 = j*16 + 8*(lane & 1);
                if (xx < W && n0 < a.cout)
                    *reinterpret_cast<bf16x8*>(
                        a.y + size_t((zz*H + yy)*W + xx)*a.ldy + n0) =
                        *reinterpret_cast<const bf16x8*>(tb + lane*8);
            }
            __syncthreads();
        }
    }
}

// Weight gradient dW[n][k] += sum over voxels of dZ[m][n] * X[tap(m)][c] for
// one 16 x 16 tile of (n, k) and one chunk of voxels per wavefront; the
// chunk partial sums are added into the FP32 gradient atomically.
// A = dZ^T (16 n x 32 voxels), B = X_col (32 voxels x 16 k).
__global__ __launch_bounds__(64) void k_conv_wgrad(WgradArgs a)
{
    const int lane = threadIdx.x;
    const int row = lane & 15;
    const int kq = (lane >> 4)*8;
    const int n = blockIdx.y*16 + row;        // A row
    const int kcol = blockIdx.x*16 + row;     // B column
    const int M = a.g.D*a.g.H*a.g.W;
    const int K = a.taps*a.cin;
    const int tcol = kcol < K ? kcol/a.cin : 0;
    const int ccol = kcol < K ? kcol - tcol*a.cin : 0;
    const int mlo = blockIdx.z*a.chunk;
    const int mhi = min(M, mlo + a.chunk);
    floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int mb = mlo; mb < mhi; mb += 32)
    {
        bf16x8 av, bv;
#pragma unroll
        for (int i = 0; i < 8; i++)
        {
            const int m = mb + kq + i;
            bf16 dz = bf16(0.0f), xv = bf16(0.0f);
            if (m < mhi)
            {
                if (n < a.cout) dz = a.dz[size_t(m)*a.lddz + n];
                if (kcol < K)
                {
                    int v = m;
                    if (a.taps != 1)
                    {
                        const int x = m % a.g.W, y = (m/a.g.W) % a.g.H;
                        const int z = m/(a.g.W*a.g.H);
                        v = tap_voxel(a.g, z, y, x, tcol);
                    }
                    if (v >= 0) xv = a.x[size_t(v)*a.ldx + ccol];
                }
            }
            av[i] = dz;
            bv[i] = xv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    // D[n][k]: lane holds column k = blockIdx.x*16 + row, rows n = 4*(lane/16)+r
#pragma unroll
    for (int r = 0; r < 4; r++)
    {
        const int nn = blockIdx.y*16 + 4*(lane >> 4) + r;
        if (nn < a.cout && kcol < K)
            atomicAdd(a.dw + size_t(nn)*a.kp + kcol, acc[r]);
    }
}

// W'[c][t'*Cout + n] = W[n][(26 - t')*Cin + c]: the weights of the data
// gradient, a forward convolution of dZ with flipped taps and swapped
// channels.
__global__ void k_flip_weights(const bf16* w, int cin, int cout, int kp,
                               int taps, bf16* wt, int kpt)
{
    const int i = blockIdx.x*blockDim.x + threadIdx.x;
    const int total = cin*kpt;
    if (i >= total) return;
    const int c = i/kpt, k = i - c*kpt;
    bf16 v = bf16(0.0f);
    if (k < taps*cout)
    {
        const int t = k/cout, n = k - t*cout;
        const int tsrc = taps == 1 ? 0 : 26 - t;
        v = w[size_t(n)*kp + tsrc*cin + c];
    }
    wt[i] = v;
}

// dZ = dY where Y > 0 (ReLU), in place on a channel slice.
__global__ void k_relu_bwd(bf16* dy, int lddy, const bf16* y, int ldy, int M,
                           int C)
{
    const long i = long(blockIdx.x)*blockDim.x + threadIdx.x;
    if (i >= long(M)*C) return;
    const int m = int(i/C), c = int(i - long(m)*C);
    if (!(bf2f(y[size_t(m)*ldy + c]) > 0.0f))
        dy[size_t(m)*lddy + c] = bf16(0.0f);
}

// 2x2x2 average pooling of a channel slice (fine grid g) into a coarse one.
__global__ void k_pool(const bf16* x, int ldx, Grid3 g, int C, bf16* y,
                       int ldy)
{
    const int Dc = g.D/2, Hc = g.H/2, Wc = g.W/2;
    const long i = long(blockIdx.x)*blockDim.x + threadIdx.x;
    if (i >= long(Dc)*Hc*Wc*C) return;
    const int c = int(i % C);
    const int mc = int(i/C);
    const int xc = mc % Wc, yc = (mc/Wc) % Hc, zc = mc/(Wc*Hc);
    float s = 0.0f;
    for (int d = 0; d < 8; d++)
    {
        const int zf = 2*zc + (d >> 2), yf = 2*yc + ((d >> 1) & 1),
                  xf = 2*xc + (d & 1);
        s += bf2f(x[size_t((zf*g.H + yf)*g.W + xf)*ldx + c]);
    }
    y[size_t(mc)*ldy + c] = bf16(0.125f*s);
}

// Gradient of the pooling: every fine voxel receives dY/8 of its parent,
// added to dX (which already holds the skip-connection gradient).
__global__ void k_pool_bwd(const bf16* dy, int lddy, Grid3 g, int C, bf16* dx,
                           int lddx)
{
    const long i = long(blockIdx.x)*blockDim.x + threadIdx.x;
    if (i >= long(g.D)*g.H*g.W*C) return;
    const int c = int(i % C);
    const int m = int(i/C);
    const int x = m % g.W, y = (m/g.W) % g.H, z = m/(g.W*g.H);
    const int mc = ((z/2)*(g.H/2) + y/2)*(g.W/2) + x/2;
    const float v = bf2f(dx[size_t(m)*lddx + c]) +
                    0.125f*bf2f(dy[size_t(mc)*lddy + c]);
    dx[size_t(m)*lddx + c] = bf16(v);
}

// Nearest-neighbour upsampling by 2 of a coarse channel slice into a fine
// one (grid g is the fine grid).
__global__ void k_up(const bf16* x, int ldx, Grid3 g, int C, bf16* y, int ldy)
{
    const long i = long(blockIdx.x)*blockDim.x + threadIdx.x;
    if (i >= long(g.D)*g.H*g.W*C) return;
    const int c = int(i % C);
    const int m = int(i/C);
    const int xx = m % g.W, yy = (m/g.W) % g.H, zz = m/(g.W*g.H);
    const int mc = ((zz/2)*(g.H/2) + yy/2)*(g.W/2) + xx/2;
    y[size_t(m)*ldy + c] = x[size_t(mc)*ldx + c];
}

// k_up and k_pool on 8 channels per thread (16-byte loads and stores;
// C, the row strides and the slice bases multiples of 8 channels), with
// the same per-channel arithmetic.
__global__ void k_up8(const bf16* x, int ldx, Grid3 g, int C, bf16* y,
                      int ldy)
{
    const int C8 = C/8;
    const long i = long(blockIdx.x)*blockDim.x + threadIdx.x;
    if (i >= long(g.D)*g.H*g.W*C8) return;
    const int c = int(i % C8)*8;
    const int m = int(i/C8);
    const int xx = m % g.W, yy = (m/g.W) % g.H, zz = m/(g.W*g.H);
    const int mc = ((zz/2)*(g.H/2) + yy/2)*(g.W/2) + xx/2;
    *reinterpret_cast<bf16x8*>(y + size_t(m)*ldy + c) =
        *reinterpret_cast<const bf16x8*>(x + size_t(mc)*ldx + c);
}

__global__ void k_pool8(const bf16* x, int ldx, Grid3 g, int C, bf16* y,
                        int ldy)
{
    const int Dc = g.D/2, Hc = g.H/2, Wc = g.W/2;
    const int C8 = C/8;
    const long i = long(blockIdx.x)*blockDim.x + threadIdx.x;
    if (i >= long(Dc)*Hc*Wc*C8) return;
    const int c = int(i % C8)*8;
    const int mc = int(i/C8);
    const int xc = mc % Wc, yc = (mc/Wc) % Hc, zc = mc/(Wc*Hc);
    float s[8];
#pragma unroll
    for (int q = 0; q < 8; q++) s[q] = 0.0f;
    for (int d = 0; d < 8; d++)
    {
        const int zf = 2*zc + (d >> 2), yf = 2*yc + ((d >> 1) & 1),
                  xf = 2*xc + (d & 1);
        const bf16x8 v = *reinterpret_cast<const bf16x8*>(
            x + size_t((zf*g.H + yf)*g.W + xf)*ldx + c);
#pragma unroll
        for (int q = 0; q < 8; q++) s[q] += float(v[q]);
    }
    bf16x8 o;
#pragma unroll
    for (int q = 0; q < 8; q++) o[q] = bf16(0.125f*s[q]);
    *reinterpret_cast<bf16x8*>(y + size_t(mc)*ldy + c) = o;
}

// Gradient of the upsampling: the coarse voxel sums its 8 children.
__global__ void k_up_bwd(const bf16* dy, int lddy, Grid3 g, int C, bf16* dx,
                         int lddx)
{
    const int Dc = g.D/2, Hc = g.H/2, Wc = g.W/2;
    const long i = long(blockIdx.x)*blockDim.x + threadIdx.x;
    if (i >= long(Dc)*Hc*Wc*C) return;
    const int c = int(i % C);
    const int mc = int(i/C);
    const int xc = mc % Wc, yc = (mc/Wc) % Hc, zc = mc/(Wc*Hc);
    float s = 0.0f;
    for (int d = 0; d < 8; d++)
    {
        const int zf = 2*zc + (d >> 2), yf = 2*yc + ((d >> 1) & 1),
                  xf = 2*xc + (d & 1);
        s += bf2f(dy[size_t((zf*g.H + yf)*g.W + xf)*lddy + c]);
    }
    dx[size_t(mc)*lddx + c] = bf16(s);
}

// Output projection (1x1x1, 16 channels -> 1, FP32; w[16] is the bias)
// and its gradients.
__global__ void k_head_fwd(const bf16* x, int ldx, int M, const float* w,
                           float* out)
{
    const int m = blockIdx.x*blockDim.x + threadIdx.x;
    if (m >= M) return;
    float s = w[16];
#pragma unroll
    for (int c = 0; c < 16; c++) s += w[c]*bf2f(x[size_t(m)*ldx + c]);
    out[m] = s;
}

// dX[m][c] = g[m] * w[c]; per-block partial sums of g[m]*X[m][c] (16
// channels) and of g[m] for the weight and bias gradients.
__global__ __launch_bounds__(256) void k_head_bwd(const bf16* x, int ldx,
                                                  int M, const float* w,
                                                  const float* gout,
                                                  bf16* dx, int lddx,
                                                  float* partial)
{
    __shared__ float sm[17][256];
    const int m = blockIdx.x*256 + threadIdx.x;
    float gw[17];
#pragma unroll
    for (int c = 0; c < 17; c++) gw[c] = 0.0f;
    if (m < M)
    {
        const float g = gout[m];
#pragma unroll
        for (int c = 0; c < 16; c++)
        {
            gw[c] = g*bf2f(x[size_t(m)*ldx + c]);
            dx[size_t(m)*lddx + c] = bf16(g*w[c]);
        }
        gw[16] = g;
    }
#pragma unroll
    for (int c = 0; c < 17; c++) sm[c][threadIdx.x] = gw[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (int(threadIdx.x) < s)
            for (int c = 0; c < 17; c++)
                sm[c][threadIdx.x] += sm[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 17) partial[size_t(blockIdx.x)*17 + threadIdx.x] =
        sm[threadIdx.x][0];
}

// Mean-removed squared error: loss partials and the output gradient
// g = 2 (p - t - mean(p - t)) / M, with mean(p - t) = shift.
__global__ __launch_bounds__(256) void k_mse(const float* p, const float* t,
                                             int M, const float* shift,
                                             float* g, float* partial)
{
    __shared__ float sm[256];
    const int m = blockIdx.x*256 + threadIdx.x;
    float e = 0.0f;
    if (m < M)
    {
        const float d = p[m] - t[m] - shift[0];
        g[m] = 2.0f*d/float(M);
        e = d*d;
    }
    sm[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (int(threadIdx.x) < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

// Per-block partial sums of a - b (FP32), for the mean shift of the loss.
__global__ __launch_bounds__(256) void k_diff_sum(const float* a,
                                                  const float* b, int M,
                                                  float* partial)
{
    __shared__ float sm[256];
    const int m = blockIdx.x*256 + threadIdx.x;
    sm[threadIdx.x] = m < M ? a[m] - b[m] : 0.0f;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (int(threadIdx.x) < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

// out[0..cols) = scale * column sums of an rows x cols FP32 array (one
// thread per column, fixed order).
__global__ void k_col_sum(const float* in, int rows, int cols, float scale,
                          float* out)
{
    const int c = blockIdx.x*blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int r = 0; r < rows; r++) s += in[size_t(r)*cols + c];
    out[c] = scale*s;
}

// Adam on FP32 master weights, refreshing the BF16 copy.
__global__ void k_adam(float* w, float* mo, float* ve, const float* g, int n,
                       float lr, float b1, float b2, float eps, float c1,
                       float c2, bf16* wb)
{
    const int i = blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float m = b1*mo[i] + (1.0f - b1)*gi;
    const float v = b2*ve[i] + (1.0f - b2)*gi*gi;
    mo[i] = m;
    ve[i] = v;
    const float wi = w[i] - lr*(m/c1)/(sqrtf(v/c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = bf16(wi);
}

__global__ void k_to_bf16(const float* a, bf16* b, long n)
{
    const long i = long(blockIdx.x)*blockDim.x + threadIdx.x;
    if (i < n) b[i] = bf16(a[i]);
}

// Network input from the solver's FP64 colour-split level-0 fields is built
// on the host side of this module (input channels: scaled right-hand side,
// scaled log coefficient, zero padding to 8).
__global__ void k_pack_input(const float* f, const float* lam, int M,
                             bf16* x)
{
    const int m = blockIdx.x*blockDim.x + threadIdx.x;
    if (m >= M) return;
    x[size_t(m)*8 + 0] = bf16(f[m]);
    x[size_t(m)*8 + 1] = bf16(lam[m]);
#pragma unroll
    for (int c = 2; c < 8; c++) x[size_t(m)*8 + c] = bf16(0.0f);
}

inline unsigned nblk(long n, int b) { return unsigned((n + b - 1)/b); }

}  // namespace

namespace {

template <int NT, int MT, int CIN>
void launch_conv_lds(const ConvArgs& a, hipStream_t s)
{
    using Tl = ConvTile<MT>;
    const int ch = CIN > 0 ? CIN : 32;
    const int vs = ch + 8;
    const size_t shm = size_t(Tl::HV)*vs*sizeof(bf16);
    const unsigned nb = unsigned((a.g.W + 15)/16)*
                        unsigned((a.g.H + Tl::TY - 1)/Tl::TY)*
                        unsigned((a.g.D + Tl::TZ - 1)/Tl::TZ);
    hipLaunchKernelGGL((k_conv_fwd_lds<NT, MT, CIN>), dim3(nb), dim3(256), shm,
                       s, a);
}

template <int NT, int CIN>
void conv_lds_mt(const ConvArgs& a, hipStream_t s)
{
    constexpr int MT0 = NT <= 4 ? 4 : (NT <= 8 ? 2 : 1);
    constexpr int ch = CIN > 0 ? CIN : 32;
    constexpr bool fit4 = ConvTile<4>::HV*(ch + 8)*2 <= 65536;
    constexpr bool fit2 = ConvTile<2>::HV*(ch + 8)*2 <= 65536;
    constexpr int MT = (MT0 >= 4 && fit4) ? 4 : ((MT0 >= 2 && fit2) ? 2 : 1);
    launch_conv_lds<NT, MT, CIN>(a, s);
}

// Output tiles per wavefront: at most 16 accumulator tiles (NT x MT), and
// a staged input tile within 64 KB of LDS (measured: Cin = 48 with four M
// tiles per wavefront in 80 KB was slower than two in 48 KB).
template <int NT>
void conv_lds(const ConvArgs& a, hipStream_t s)
{
    switch (a.cin)
    {
    case 8: conv_lds_mt<NT, 8>(a, s); return;
    case 16: conv_lds_mt<NT, 16>(a, s); return;
    case 32: conv_lds_mt<NT, 32>(a, s); return;
    case 48: conv_lds_mt<NT, 48>(a, s); return;
    default: conv_lds_mt<NT, 0>(a, s); return;
    }
}

}  // namespace

void conv_fwd(const ConvArgs& a, hipStream_t s)
{
    const int M = a.g.D*a.g.H*a.g.W;
    const dim3 g(nblk(M, 64));
    if (a.cout % 16 || a.cin % 8 || a.ldx % 8 || a.kp % 32)
        throw UnetError("convolution shape not supported");
    if (a.taps == 27 && (a.cin == 8 || a.cin == 16 || a.cin == 48 ||
                         a.cin % 32 == 0))
    {
        switch (a.cout/16)
        {
        case 1: conv_lds<1>(a, s); return;
        case 2: conv_lds<2>(a, s); return;
        case 3: conv_lds<3>(a, s); return;
        case 4: conv_lds<4>(a, s); return;
        case 6: conv_lds<6>(a, s); return;
        case 8: conv_lds<8>(a, s); return;
        case 12: conv_lds<12>(a, s); return;
        default: throw UnetError("unsupported output channel count");
        }
    }
    switch (a.cout/16)
    {
    case 1: hipLaunchKernelGGL((k_conv_fwd<1>), g, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_conv_fwd<2>), g, dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_conv_fwd<3>), g, dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_conv_fwd<4>), g, dim3(256), 0, s, a); break;
    case 6: hipLaunchKernelGGL((k_conv_fwd<6>), g, dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_conv_fwd<8>), g, dim3(256), 0, s, a); break;
    case 12: hipLaunchKernelGGL((k_conv_fwd<12>), g, dim3(256), 0, s, a); break;
    default: throw UnetError("unsupported output channel count");
    }
}

void conv_wgrad(const WgradArgs& a, hipStream_t s)
{
    const int M = a.g.D*a.g.H*a.g.W;
    const int K = a.taps*a.cin;
    const dim3 g(nblk(K, 16), nblk(a.cout, 16), nblk(M, a.chunk));
    hipLaunchKernelGGL(k_conv_wgrad, g, dim3(64), 0, s, a);
}

void flip_weights(const bf16* w, int cin, int cout, int kp, int taps, bf16* wt,
                  int kpt, hipStream_t s)
{
    hipLaunchKernelGGL(k_flip_weights, dim3(nblk(long(cin)*kpt, 256)),
                       dim3(256), 0, s, w, cin, cout, kp, taps, wt, kpt);
}

void relu_bwd(bf16* dy, int lddy, const bf16* y, int ldy, int M, int C,
              hipStream_t s)
{
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(long(M)*C, 256)), dim3(256), 0, s,
                       dy, lddy, y, ldy, M, C);
}

void pool(const bf16* x, int ldx, Grid3 g, int C, bf16* y, int ldy,
          hipStream_t s)
{
    const long n = long(g.D/2)*(g.H/2)*(g.W/2)*C;
    if (C % 8 == 0 && ldx % 8 == 0 && ldy % 8 == 0 &&
        reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
        reinterpret_cast<uintptr_t>(y) % 16 == 0)
        hipLaunchKernelGGL(k_pool8, dim3(nblk(n/8, 256)), dim3(256), 0, s, x,
                           ldx, g, C, y, ldy);
    else
        hipLaunchKernelGGL(k_pool, dim3(nblk(n, 256)), dim3(256), 0, s, x,
                           ldx, g, C, y, ldy);
}

void pool_bwd(const bf16* dy, int lddy, Grid3 g, int C, bf16* dx, int lddx,
              hipStream_t s)
{
    const long n = long(g.D)*g.H*g.W*C;
    hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(n, 256)), dim3(256), 0, s, dy,
                       lddy, g, C, dx, lddx);
}

void up(const bf16* x, int ldx, Grid3 g, int C, bf16* y, int ldy,
        hipStream_t s)
{
    const long n = long(g.D)*g.H*g.W*C;
    if (C % 8 == 0 && ldx % 8 == 0 && ldy % 8 == 0 &&
        reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
        reinterpret_cast<uintptr_t>(y) % 16 == 0)
        hipLaunchKernelGGL(k_up8, dim3(nblk(n/8, 256)), dim3(256), 0, s, x, ldx,
                           g, C, y, ldy);
    else
        hipLaunchKernelGGL(k_up, dim3(nblk(n, 256)), dim3(256), 0, s, x, ldx, g,
                           C, y, ldy);
}

void up_bwd(const bf16* dy, int lddy, Grid3 g, int C, bf16* dx, int lddx,
            hipStream_t s)
{
    const long n = long(g.D/2)*(g.H/2)*(g.W/2)*C;
    hipLaunchKernelGGL(k_up_bwd, dim3(nblk(n, 256)), dim3(256), 0, s, dy, lddy,
                       g, C, dx, lddx);
}

void head_fwd(const bf16* x, int ldx, int M, const float* w, float* out,
              hipStream_t s)
{
    hipLaunchKernelGGL(k_head_fwd, dim3(nblk(M, 256)), dim3(256), 0, s, x, ldx,
                       M, w, out);
}

int head_bwd(const bf16* x, int ldx, int M, const float* w, const float* gout,
             bf16* dx, int lddx, float* partial, hipStream_t s)
{
    const unsigned nb = nblk(M, 256);
    hipLaunchKernelGGL(k_head_bwd, dim3(nb), dim3(256), 0, s, x, ldx, M, w,
                       gout, dx, lddx, partial);
    return int(nb);
}

int mse(const float* p, const float* t, int M, const float* shift, float* g,
        float* partial, hipStream_t s)
{
    const unsigned nb = nblk(M, 256);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, s, p, t, M, shift, g,
                       partial);
    return int(nb);
}

int diff_sum(const float* a, const float* b, int M, float* partial,
             hipStream_t s)
{
    const unsigned nb = nblk(M, 256);
    hipLaunchKernelGGL(k_diff_sum, dim3(nb), dim3(256), 0, s, a, b, M, partial);
    return int(nb);
}

void col_sum(const float* in, int rows, int cols, float scale, float* out,
             hipStream_t s)
{
    hipLaunchKernelGGL(k_col_sum, dim3(nblk(cols, 64)), dim3(64), 0, s, in,
                       rows, cols, scale, out);
}

void adam(float* w, float* m, float* v, const float* g, int n, float lr,
          float b1, float b2, float eps, int step, bf16* wb, hipStream_t s)
{
    const float c1 = 1.0f - powf(b1, float(step));
    const float c2 = 1.0f - powf(b2, float(step));
    hipLaunchKernelGGL(k_adam, dim3(nblk(n, 256)), dim3(256), 0, s, w, m, v, g,
                       n, lr, b1, b2, eps, c1, c2, wb);
}

void to_bf16(const float* a, bf16* b, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n, 256)), dim3(256), 0, s, a, b, n);
}

void pack_input(const float* f, const float* lam, int M, bf16* x,
                hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_input, dim3(nblk(M, 256)), dim3(256), 0, s, f,
                       lam, M, x);
}

}  // namespace unet
}  // namespace bmg
