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

void conv_fwd(const ConvArgs& a, hipStream_t s)
{
    const int M = a.g.D*a.g.H*a.g.W;
    const dim3 g(nblk(M, 64));
    if (a.cout % 16 || a.cin % 8 || a.ldx % 8 || a.kp % 32)
        throw UnetError("convolution shape not supported");
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
    hipLaunchKernelGGL(k_pool, dim3(nblk(n, 256)), dim3(256), 0, s, x, ldx, g,
                       C, y, ldy);
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
    hipLaunchKernelGGL(k_up, dim3(nblk(n, 256)), dim3(256), 0, s, x, ldx, g, C,
                       y, ldy);
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
