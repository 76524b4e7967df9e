/*
 * ppf_unet.h -- learned K-step jump for the phase-field state: a 3-level
 * 3-D convolutional U-Net on MI355X (part of libppf.so).
 *
 * No counterpart in the reference, which only steps explicitly
 * (phasefield.c:178-258).  The network maps (etaS, etaL, c) at step t to the
 * state at step t+K:
 *
 *   level 1 (N^3):    conv3 3->B, conv3 B->B                 (skip)
 *   level 2 (N^3/8):  avgpool2, conv3 B->2B, conv3 2B->2B   (skip)
 *   level 3 (N^3/64): avgpool2, conv3 2B->4B, conv3 4B->4B
 *   decoder:          nearest x2, concat skip, conv3 6B->2B, conv3 2B->2B,
 *                     nearest x2, concat skip, conv3 3B->B,  conv3 B->B,
 *   head:             1x1x1 conv B->3; output = input + head
 *
 * with B = 32 base channels, 3x3x3 kernels with periodic padding, leaky
 * ReLU (slope 0.1) after every 3x3x3 conv.  Weights and activations are
 * BF16 on the MFMA units with FP32 accumulation; master weights, gradients
 * and the Adam moments are FP32.  Grids must have every dimension a
 * multiple of 4 (two poolings).
 *
 * Parameters, flat FP32, layer by layer: W[27][Cout][Cin] then b[Cout] for
 * the ten 3x3x3 convolutions (tap = (dz+1)*9 + (dy+1)*3 + (dx+1): the weight
 * of input site (z+dz, y+dy, x+dx)), then the head's W[3][B] and b[3].
 *
 * Fields crossing this ABI are float32, [batch][nz][ny][nx][3] with the
 * channels etaS, etaL, c last.  Every call returns 0 or a negative PPF_E*.
 */
#ifndef PPF_UNET_H
#define PPF_UNET_H

#include <stdint.h>

#include "ppf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ppf_unet ppf_unet;

/* base: channels at level 1 (32 or 64); seed: the seeded uniform init
 * (He, for leaky ReLU); the head starts at zero, so an untrained network
 * returns its input. */
int ppf_unet_create(ppf_unet **out, int base, int device, uint64_t seed);
int ppf_unet_destroy(ppf_unet *net);
int ppf_unet_nparams(ppf_unet *net, int64_t *n);
int ppf_unet_get_params(ppf_unet *net, float *params);
int ppf_unet_set_params(ppf_unet *net, const float *params);

/* y = net(x) for a batch of grids (host arrays). */
int ppf_unet_forward(ppf_unet *net, int batch, int nz, int ny, int nx, const float *x, float *y);

/* loss = mean((net(x) - y)^2) over batch, sites and 3 channels, and its
 * gradient with respect to every parameter (flat, as get_params); no
 * update.  grads may be NULL. */
int ppf_unet_loss_grad(ppf_unet *net, int batch, int nz, int ny, int nx, const float *x,
                       const float *y, float *loss, float *grads);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that loss. */
int ppf_unet_train_step(ppf_unet *net, int batch, int nz, int ny, int nx, const float *x,
                        const float *y, float lr, float *loss);

/* Apply the network to a stepping context's current state in place (a
 * jump of K steps).  c is then shifted uniformly so that sum(c) equals its
 * value before the jump (FP64 sums), conserving solute exactly.  Queued in
 * stream order with the context's work (no host synchronisation): steps
 * issued on the context before the call run first, and the context's later
 * calls see the jumped state. */
int ppf_unet_apply(ppf_unet *net, ppf_ctx *ctx);

/* Milliseconds of the last forward pass, measured with HIP events, and
 * the MFMA work its kernels executed, in FLOP: 2*27*Cin*Cout per site per
 * 3x3x3 conv, the first layer's K padded to 160, and 8 instead of 27 taps
 * on the upsampled channels of the polyphase decoder layers (kind 6). */
int ppf_unet_last_forward(ppf_unet *net, double *ms, double *flop);

/* Convolution kernel:
 *   2 row-blocked: a wave owns 16 x-sites of 8 rows and 32 or 64
 *     output channels and reuses each loaded input row for the three dy
 *     taps; no LDS; the first layer runs as one GEMM over its 81 real (tap,
 *     channel) pairs instead of 27 taps x 32 padded channels;
 *     at inference the two decoder convolutions read their upsampled
 *     channels straight from the coarse tensor (no upsampled copy);
 *   3, 4 the same with other tiles: 8 rows x 32 channels everywhere (3), or
 *     4 rows x 64 channels where Cout allows (4), for tuning;
 *   5 kind 2 with the upsampled channels materialized, as in training (the
 *     same sums as kind 2, for cross-checks);
 *   6 (default) kind 2 with the two decoder convolutions in polyphase form at
 *     inference: for each output parity, the 3x3x3 convolution of the
 *     nearest-upsampled channels is a 2x2x2 convolution of the coarse
 *     tensor with summed weights (8 taps instead of 27); the summed weights
 *     are rounded to BF16 once, so kind 6 agrees with kind 2 to BF16 weight
 *     rounding, not bitwise; training runs as kind 2; the last decoder
 *     convolution of the base-32 network runs z-marching, with its fine and
 *     coarse inputs staged in LDS and its K dimension split over two waves
 *     per SIMD;
 *   7 kind 6 with the row-blocked polyphase kernel for both decoder
 *     convolutions (kind 6's products; in the last decoder convolution
 *     summed in one instead of two partial sums, so the two agree to FP32
 *     rounding before the BF16 output rounding);
 *   1 LDS-tiled: a 16 x 8 brick's halo staged in LDS;
 *   0 direct: every operand straight from global memory.
 * Kinds 2-5 add the same (tap, 32-channel) products in the same order:
 * their forwards are bitwise equal.  Kinds 0 and 1 sum in other orders and
 * agree with them to FP32 rounding. */
int ppf_unet_set_conv_kernel(ppf_unet *net, int kind);

const char *ppf_unet_last_error(ppf_unet *net);

#ifdef __cplusplus
}
#endif
#endif
