// Internal declarations of the learned preconditioner (3-D U-Net, BF16 on
// the gfx950 matrix cores).
#pragma once

#include <hip/hip_runtime.h>
#include <stdexcept>
#include <string>

struct bmg_unet;

namespace bmg {
namespace unet {

typedef __bf16 bf16;

struct UnetError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

struct Grid3 {
    int D, H, W;   // z, y, x extents; voxel m = (z*H + y)*W + x
};

// One convolution y = act(conv(x, w) + bias) on channel slices.
struct ConvArgs {
    Grid3 g;
    const bf16* x;    // input slice, row stride ldx, cin channels (cin % 8 == 0)
    int ldx;
    int cin;
    const bf16* w;    // [cout][kp], k = t*cin + c, zero beyond taps*cin
    int kp;           // taps*cin rounded up to 32
    int taps;         // 27 (3x3x3) or 1 (1x1x1)
    int cout;         // a multiple of 16 up to 192
    const float* bias;  // cout values or null
    int relu;
    bf16* y;          // output slice, row stride ldy
    int ldy;
};

// dw[n][k] += sum_m dz[m][n] * x[tap_k(m)][c_k] (FP32 accumulation).
struct WgradArgs {
    Grid3 g;
    const bf16* x;
    int ldx;
    int cin;
    const bf16* dz;
    int lddz;
    int cout;
    int taps;
    int kp;
    float* dw;
    int chunk;        // voxels per workgroup
};

inline int round32(int k) { return (k + 31)/32*32; }

void conv_fwd(const ConvArgs& a, hipStream_t s);
void conv_wgrad(const WgradArgs& a, hipStream_t s);
// weights of the data gradient: wt[c][t'*cout + n] = w[n][(26-t')*cin + c]
void flip_weights(const bf16* w, int cin, int cout, int kp, int taps, bf16* wt,
                  int kpt, hipStream_t s);
void relu_bwd(bf16* dy, int lddy, const bf16* y, int ldy, int M, int C,
              hipStream_t s);
void pool(const bf16* x, int ldx, Grid3 fine, int C, bf16* y, int ldy,
          hipStream_t s);
void pool_bwd(const bf16* dy, int lddy, Grid3 fine, int C, bf16* dx, int lddx,
              hipStream_t s);
void up(const bf16* x, int ldx, Grid3 fine, int C, bf16* y, int ldy,
        hipStream_t s);
void up_bwd(const bf16* dy, int lddy, Grid3 fine, int C, bf16* dx, int lddx,
            hipStream_t s);
// w[0..15] weights, w[16] bias
void head_fwd(const bf16* x, int ldx, int M, const float* w, float* out,
              hipStream_t s);
int head_bwd(const bf16* x, int ldx, int M, const float* w, const float* gout,
             bf16* dx, int lddx, float* partial, hipStream_t s);
int mse(const float* p, const float* t, int M, const float* shift, float* g,
        float* partial, hipStream_t s);
int diff_sum(const float* a, const float* b, int M, float* partial,
             hipStream_t s);
void col_sum(const float* in, int rows, int cols, float scale, float* out,
             hipStream_t s);
void adam(float* w, float* m, float* v, const float* g, int n, float lr,
          float b1, float b2, float eps, int step, bf16* wb, hipStream_t s);
void to_bf16(const float* a, bf16* b, long n, hipStream_t s);
void pack_input(const float* f, const float* lam, int M, bf16* x,
                hipStream_t s);

// Solver-facing calls on a network (bmg_unet of bmg_unet.h): allocate the
// activations of a grid (not allowed during graph capture), and run the
// forward pass on a given stream (capturable once prepared).
void net_prepare(struct ::bmg_unet* net, int D, int H, int W);
void net_forward(struct ::bmg_unet* net, int D, int H, int W, const float* f,
                 const float* c, float* out, hipStream_t s);

}  // namespace unet
}  // namespace bmg
