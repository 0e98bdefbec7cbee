/*
 * dsmc_surrogate.h - C-ABI of the learned warm start of the MI355X (gfx950) DSMC pipeline.
 *
 * The reference has no counterpart: it always starts from the freestream fill of
 * InitialConditions (src/Initialization.cpp:290-313). This library adds a small convolutional
 * encoder-decoder that maps a case's parameters (Mach, angle of attack, density, wall model,
 * accommodation, wall temperature) and its obstacle mask to the steady fields n, ux, uy, T.
 * Its prediction seeds the particles through dsmc_init_from_fields (dsmc_c.h), from local
 * Maxwellians, so the transient before the first sampling step shrinks. The DSMC's own sampled
 * field remains the output; the freestream fill remains the fallback.
 *
 * Network: 4 levels of two 3x3 convolutions + ReLU (channels C, 2C, 4C, 8C; C = 32 gives
 * 1.95M parameters), 2x2 average pooling down, nearest upsampling up with skip concatenation,
 * a 1x1 head to 4 outputs. The 3x3 convolutions run on bf16 MFMA (forward, data gradient,
 * weight gradient) with FP32 accumulation; weights are kept in FP32 for Adam and rounded to bf16
 * for the convolutions.
 *
 * Grids: the network sees an H x W raster of the case domain (H cells along x, W along y,
 * H and W multiples of 8). Field arrays are [case][4][H][W] floats, x-major like the
 * solver's k*NY+l cell order: plane 0 n [m^-3], 1 ux, 2 uy [m/s], 3 T [K].
 *
 * Conventions as dsmc_c.h: 0 or a DSMC_E_* code; dsmc_net_last_error explains the last one.
 */
#ifndef DSMC_SURROGATE_H
#define DSMC_SURROGATE_H

#include <stddef.h>
#include <stdint.h>

#include "dsmc_c.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DSMC_NET_IN_CHANNELS 32   /* input channels as stored: 12 used, zero padded */
#define DSMC_NET_OUT_CHANNELS 4   /* normalised log(n/n_inf), ux/U_inf, uy/U_inf, log(T/T_inf) */

typedef struct dsmc_net_config {
    int32_t height, width;   /* raster (multiples of 8); 256 x 256 by default */
    int32_t base_channels;   /* C: 32 or 64 */
    int32_t max_batch;       /* largest batch of forward / training calls */
    uint64_t seed;           /* weight initialisation (He normal) */
} dsmc_net_config;

typedef struct dsmc_net dsmc_net;

void dsmc_net_config_default(dsmc_net_config* cfg);
int dsmc_net_create(const dsmc_net_config* cfg, int device, dsmc_net** out);
void dsmc_net_destroy(dsmc_net* net);
const char* dsmc_net_last_error(const dsmc_net* net);

/* Parameters, flat FP32: per layer W [cout][9][cin] then b [cout], layers in the order
 * e0a e0b e1a e1b e2a e2b e3a e3b d2a d2b d1a d1b d0a d0b, then the head Wh [4][C], bh [4]. */
int64_t dsmc_net_num_params(const dsmc_net* net);
int dsmc_net_get_weights(dsmc_net* net, float* out, size_t n);
int dsmc_net_set_weights(dsmc_net* net, const float* in, size_t n);

/* Input rasters of cases, [ncases][H][W][32] floats (exactly the bf16 values the network sees):
 * channel 0 the obstacle mask, 1-2 cell-centre coordinates in [-1, 1], 3-11 the features
 * (Mach-5)/3, AoA/30deg, log10(n)-17, wall one-hot (diffuse, specular, CLL), effective acc_n
 * and acc_t (diffuse 1, specular 0), T_wall/T_inf - 1.5. */
int dsmc_net_encode(dsmc_net* net, const dsmc_params* cases, int ncases, float* inputs);
/* Training targets from physical fields [n][4][H][W]: targets [n][H][W][4] (the normalised
 * outputs) and per-pixel loss weights [n][H][W] (1 / (4 * gas pixels of the case) on pixels
 * outside the obstacle with n > 0 and T > 0, else 0). Host only; needs no device. */
int dsmc_net_normalize(const dsmc_net_config* cfg, const dsmc_params* cases, int n,
                       const float* fields, float* targets, float* weights);

/* Raw forward on given inputs [B][H][W][32] (rounded to bf16): outputs [B][H][W][4]. */
int dsmc_net_forward_raw(dsmc_net* net, const float* inputs, int batch, float* outputs);
/* After a forward call: the bf16 activation of layer `layer` (0..13 in the order above,
 * 14..16 the three pooled tensors; after dsmc_net_loss_grad_raw, 17..30 the gradient at the
 * output of convolution 0..13 after its ReLU mask) as floats [B][h][w][channels]; *count
 * receives its size. */
int dsmc_net_get_activation(dsmc_net* net, int layer, float* out, size_t cap, size_t* count);
/* Loss (mean over the batch of each sample's weighted squared error) and its gradient with
 * respect to every parameter (flat, as dsmc_net_get_weights), on raw inputs, targets
 * [B][H][W][4] and weights [B][H][W]. */
int dsmc_net_loss_grad_raw(dsmc_net* net, const float* inputs, const float* targets,
                           const float* weights, int batch, float* loss, float* grad, size_t n);

/* Training set kept in device memory: case parameters and their physical steady fields
 * [n][4][H][W]. Replaces any previous set. */
int dsmc_net_set_dataset(dsmc_net* net, const dsmc_params* cases, const float* fields, int n);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over `epochs` shuffled passes in batches of `batch`
 * (the last partial batch of an epoch is dropped), learning rate lr with a cosine decay to
 * lr/20. loss_history[step] receives each step's batch loss (may be null). Continues the
 * optimiser state of earlier calls. */
int dsmc_net_train(dsmc_net* net, int epochs, int batch, float lr, uint64_t shuffle_seed,
                   float* loss_history, size_t history_cap);
/* Mean loss over a set of cases and fields (no update). */
int dsmc_net_evaluate(dsmc_net* net, const dsmc_params* cases, const float* fields, int n,
                      float* loss);
/* Predicted physical fields [ncases][4][H][W] for dsmc_init_from_fields: n = n_inf
 * exp(out0), ux = U_inf out1, uy = U_inf out2, T = T_inf exp(out3); n = 0 inside the obstacle. */
int dsmc_net_predict(dsmc_net* net, const dsmc_params* cases, int ncases, float* fields);

/* Milliseconds of the last dsmc_net_train call, and its MFMA work in FLOPs (3x3 convolutions
 * only: forward, data gradient and weight gradient). */
int dsmc_net_last_train_stats(dsmc_net* net, double* ms, double* conv_flops);

#ifdef __cplusplus
}
#endif

#endif /* DSMC_SURROGATE_H */
