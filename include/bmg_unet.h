/*
 * bmg_unet.h -- C ABI of the learned preconditioner of the MI355X multigrid
 * solver: a 4-level 3-D convolutional U-Net (channels 16/32/64/128, 3x3x3
 * convolutions, BF16 weights and activations, FP32 accumulation) with its
 * own forward and backward kernels on the gfx950 matrix cores, trained on
 * the device on (right-hand side, solution) pairs the multigrid generates,
 * and used by the solver as initial guess and preconditioner of flexible
 * CG (BMG_KRYLOV_FCG_LEARNED) with a fall-back to multigrid cycles.
 *
 * None of these entry points has a counterpart in the reference: Briscola's
 * solvers (src/briscolaFiniteVolume/solvers) have no learned component.
 *
 * Network input (device FP32, D x H x W each, voxel m = (z*H + y)*W + x,
 * i.e. the reference cell order with the last axis fastest):
 *   f = b / rms(b): the finite-volume right-hand side b scaled to unit RMS;
 *   c = (log10(lambda) + 1.5) / 1.5: lambda the mean of the six face
 *       coefficients of the cell (1 for the coefficient-free Laplacian).
 * Output: t = x h s / rms(b), x the solution, h the level-0 cell width along
 * the first axis (domain length / cells) and s = 10^mean(log10 lambda) the
 * geometric mean of the cell coefficients.  The finite-volume operator
 * scales with h s, so this map depends neither on the resolution nor on the
 * coefficient level.  D, H, W multiples of 8.
 * Every call returns 0 or a negative status; bmg_unet_last_error() says why.
 */
#ifndef BMG_UNET_H
#define BMG_UNET_H

#include <stdint.h>
#include "bmg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bmg_unet bmg_unet;

const char* bmg_unet_last_error(void);

/* New network with He-normal weights from a SplitMix64 stream of `seed`. */
int bmg_unet_create(bmg_unet** out, unsigned long long seed);
int bmg_unet_destroy(bmg_unet* net);

/* FP32 master parameters (all convolutions, then the head), host copies. */
long bmg_unet_num_params(const bmg_unet* net);
int bmg_unet_get_params(const bmg_unet* net, float* host);
int bmg_unet_set_params(bmg_unet* net, const float* host);
/* Gradient of the last training step (same layout as the parameters). */
int bmg_unet_get_grads(const bmg_unet* net, float* host);

/* out = U(f, c) on device buffers. */
int bmg_unet_forward(bmg_unet* net, int D, int H, int W, const float* f,
                     const float* c, float* out);

/* One Adam step (lr; betas 0.9, 0.999) on the mean-removed squared error
 * between U(f, c) and target (device buffers); *loss gets the loss before the
 * step. */
int bmg_unet_train_step(bmg_unet* net, int D, int H, int W, const float* f,
                        const float* c, const float* target, float lr,
                        float* loss);

/* Test hooks on device buffers (BF16 as uint16_t, channels-last with row
 * strides ldx/ldy/...; weights [cout][kp], k = tap*cin + c, kp = 27*cin
 * rounded up to 32, zero-filled):
 *   y = act(conv3x3x3(x, w) + bias)        (taps 27) or 1x1x1 (taps 1);
 *   dx = data gradient of conv3x3x3 for output gradient dz;
 *   dw += weight gradient (FP32 [cout][kp]), voxels in chunks of `chunk`. */
int bmg_unet_op_conv(int D, int H, int W, const uint16_t* x, int ldx, int cin,
                     const uint16_t* w, int kp, int taps, int cout,
                     const float* bias, int relu, uint16_t* y, int ldy);
int bmg_unet_op_conv_dgrad(int D, int H, int W, const uint16_t* dz, int lddz,
                           int cout, const uint16_t* w, int kp, int cin,
                           uint16_t* dx, int lddx);
int bmg_unet_op_conv_wgrad(int D, int H, int W, const uint16_t* x, int ldx,
                           int cin, const uint16_t* dz, int lddz, int cout,
                           int kp, int chunk, float* dw);

/* Attach a network to a solver context (NULL detaches).  With
 * opts.krylov = BMG_KRYLOV_FCG_LEARNED the solve starts from the network's
 * guess and preconditions with it until an iteration reduces the residual
 * by less than a factor 0.3 (drop > 0.3), then continues with multigrid
 * cycles.  The context does not own the network. */
int bmg_set_learned(bmg_ctx* ctx, bmg_unet* net);

#ifdef __cplusplus
}
#endif

#endif
