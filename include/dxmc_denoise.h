/* Dose denoiser of the MI355X transport engine (part of libdxmc_hip.so).
 *
 * No counterpart in the reference: DXMClib runs plain Monte Carlo only.  A
 * 3-D residual convolutional network maps a low-history dose map to a
 * high-history one; plain high-history Monte Carlo is its fallback and the
 * judge of its error.
 *
 * Network: 18 convolutions 3x3x3, 32 channels, stride 1, zero padding 1.
 *   layer 0      : 4 input features -> 32, ReLU
 *   layers 1..16 : 8 residual blocks  y = ReLU(x + conv(ReLU(conv(x))))
 *   layer 17     : 32 -> 1 correction of the relative dose (no activation)
 *   prediction   = input relative dose + correction.
 * Input features per voxel: relative dose D / D_ref (D_ref = mean dose over
 * voxels of material != 0), density (g/cm3), a material embedding scalar
 * (mass attenuation at 60 keV over water's), relative standard error.
 * Activations and weights are BF16 in channel-last layout; every convolution
 * is an implicit GEMM on CDNA4 BF16 matrix cores with FP32 accumulation;
 * Adam keeps FP32 master weights.  Volumes must have an x extent that is a
 * multiple of 64 (callers pad).
 */
#ifndef DXMC_DENOISE_H
#define DXMC_DENOISE_H

#include "dxmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dxg_denoiser dxg_denoiser;

enum { DXG_DENOISE_LAYERS = 18, DXG_DENOISE_CHANNELS = 32, DXG_DENOISE_FEATURES = 4 };

/* Parameters in FP32: per layer 32 x 27 x 32 weights [co][tap][ci] then 32
 * biases; padded channels (ci >= 4 in layer 0, co >= 1 in layer 17) are zero. */
uint64_t dxg_denoiser_num_params(void);

/* He-initialised network (seeded), correction layer zero (the identity map). */
int dxg_denoiser_create(dxg_denoiser** out, uint64_t seed);
void dxg_denoiser_destroy(dxg_denoiser* d);
int dxg_denoiser_get_params(const dxg_denoiser* d, float* params, uint64_t n);
int dxg_denoiser_set_params(dxg_denoiser* d, const float* params, uint64_t n);
/* Gradient of the loss of the last training step, same layout as the
 * parameters (before that step's Adam update was applied). */
int dxg_denoiser_get_grads(const dxg_denoiser* d, float* grads, uint64_t n);

/* One Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8) on a batch of patches.
 * features: batch x D x H x W x 4 (float), target / mask: batch x D x H x W.
 * Loss = sum(mask (pred - target)^2) / sum(mask), written to *loss. */
int dxg_denoiser_train_step(dxg_denoiser* d, const float* features, const float* target, const float* mask,
                            uint64_t batch, const uint64_t dims[3], float lr, float* loss);

/* Prediction (relative dose) for one volume, host arrays; D x H x W. */
int dxg_denoiser_infer(dxg_denoiser* d, const float* features, const uint64_t dims[3], float* prediction);

/* Kernel timing of the last call (ms, HIP events). */
double dxg_denoiser_last_ms(const dxg_denoiser* d);

/* ---- on-device path from a built world after dxg_transport_run: nothing of
 * the grid crosses PCIe.  D_ref is the mean dose over voxels of material != 0.
 *
 * The world's current features computed on the device, as the host computes
 * them from DoseScore; x is padded with zero features to a multiple of 64 and
 * n must be 4 * nz * ny * nx_padded. */
int dxg_denoiser_world_features(dxg_denoiser* d, const dxg_world* w, float* features, uint64_t n);

/* Denoise the world's current dose: features, inference and the denoised
 * dose in mGy (grid order, x fastest), copied to dose_mgy unless it is null.
 * dxg_denoiser_last_ms() then gives the device time of the three steps. */
int dxg_denoiser_apply(dxg_denoiser* d, const dxg_world* w, double* dose_mgy);

/* Test aid (no counterpart in the reference): the body-tile list that
 * dxg_denoiser_apply runs convolution `layer` (0..17) on — the tiles within
 * 19 - layer voxels of material != 0 — built on the device (host = 0) or by
 * the host scan (host = 1).  With row pairs (the default) an entry is
 * pair << 2 | rows to store, pair = (z * ceil(ny / 2) + y / 2) * nx_tiles + x / 64,
 * ascending.  Up to cap entries go to out; *count receives the length. */
int dxg_denoiser_debug_tiles(dxg_denoiser* d, const dxg_world* w, int host, int layer, int32_t* out, uint64_t cap,
                             uint64_t* count);

/* ---- numerics probes (tests): one convolution on host arrays.
 * x: N x D x H x W x 32, w: 32 x 27 x 32 [co][tap][ci], bias 32, y out.
 * mode 0: y = conv(x) + bias;  mode 1: dx = conv^T(dy = x) (input gradient,
 * bias ignored);  mode 2: dw[co][tap][ci] = sum_v dy[v][co] x_shift[v][ci]
 * with x = the input and `bias_or_dy` the output gradient (N x D x H x W x 32);
 * out then holds the 27648 weight gradients [co][tap][ci] and the 32 bias
 * gradients.  Mode 3: the network's first layer, y = relu(conv(x) + bias)
 * rounded to BF16, for x whose channels 4..31 are zero (the four features).
 * Modes 0, 1 and 3 write N x D x H x W x 32 floats.  Inputs are rounded
 * to BF16 first; mode 2 needs a y extent that is a multiple of 4. */
int dxg_conv3d_probe(int mode, const float* x, const float* w, const float* bias_or_dy, uint64_t n, const uint64_t dims[3],
                     float* out);

/* MFMA register-layout probe: 16x32 A and 32x16 B of small integers are
 * multiplied by one v_mfma_f32_16x16x32_bf16 under the layout the conv
 * kernels assume; d receives the 16x16 product as the kernels read it. */
int dxg_mfma_probe(const float* a, const float* b, float* d);

#ifdef __cplusplus
}
#endif

#endif /* DXMC_DENOISE_H */
