/*
 * fornos_denoise.h -- C ABI of the learned AO / bent-normal denoiser.
 *
 * No reference counterpart: the reference bakes AO and bent normals by brute
 * force only (Src/solver_ao.cpp, Src/solver_bentnormals.cpp).  The denoiser
 * is a 2-D U-Net over the baked map (3x3 convolutions + ReLU, 2x2 average
 * pooling, nearest upsampling with skip concatenation, a 1x1 residual head);
 * its convolutions run on the bf16 matrix cores with fp32 accumulation, in
 * training (forward, input and weight gradients, Adam) and in inference.
 *
 * Images are NHWC float32 in device memory, passed as plain pointers:
 *   input  [n][h][w][FORNOS_DN_IN]:  0 AO (16 samples), 1-3 bent normal
 *          (16 samples, 0 where undefined), 4-6 position, 7-9 normal,
 *          10 mapping distance, 11 texel mask (1 inside the texel set);
 *   output / target [n][h][w][FORNOS_DN_OUT]: 0 AO, 1-3 bent normal.
 * h and w must be multiples of 2^(levels-1) (and of 2).  Calls run on the
 * default stream and return when their work is complete; a handle must not
 * be used from two threads at once.
 */
#ifndef FORNOS_DENOISE_H
#define FORNOS_DENOISE_H

#include "fornos_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FORNOS_DN_IN 12
#define FORNOS_DN_OUT 4
#define FORNOS_DN_MASK_CHANNEL 11

typedef struct fornos_denoiser fornos_denoiser;

/* A U-Net with `levels` resolutions (2..6) and base_channels (a multiple of
 * 32) at the top level, doubling per level; weights He-initialised from seed. */
fornos_status fornos_denoiser_create(uint32_t levels, uint32_t base_channels, uint64_t seed,
                                     fornos_denoiser **out);
/* Parameters: per convolution layer (encoder levels top-down, then decoder
 * levels bottom-up; two 3x3 layers each) weights [cout][3][3][cin] then bias
 * [cout]; then the head's weights [4][base] and bias [4].  Layer shapes:
 * fornos_denoiser_layer. */
fornos_status fornos_denoiser_param_count(const fornos_denoiser *d, uint64_t *count);
fornos_status fornos_denoiser_layer_count(const fornos_denoiser *d, uint32_t *count);
/* c1: channels of the first input (half resolution when up != 0), c2 of the
 * skip input, cout; offset of the layer's weights in the parameter vector. */
fornos_status fornos_denoiser_layer(const fornos_denoiser *d, uint32_t i, uint32_t *c1, uint32_t *c2,
                                    uint32_t *cout, uint32_t *up, uint64_t *offset);
fornos_status fornos_denoiser_get_params(const fornos_denoiser *d, float *host, uint64_t count);
fornos_status fornos_denoiser_set_params(fornos_denoiser *d, const float *host, uint64_t count);
/* Inference; *ms = device time of the network (HIP events), may be null. */
fornos_status fornos_denoiser_forward(fornos_denoiser *d, const float *input, uint32_t n, uint32_t h, uint32_t w,
                                      float *output, float *ms);
/* Forward + masked L1 loss (mean over masked pixels of the summed absolute
 * error of the 4 outputs) + backward; the parameter gradient is kept in the
 * handle.  output may be null. */
fornos_status fornos_denoiser_backward(fornos_denoiser *d, const float *input, const float *target, uint32_t n,
                                       uint32_t h, uint32_t w, float *output, float *loss);
fornos_status fornos_denoiser_get_grads(const fornos_denoiser *d, float *host, uint64_t count);
/* One Adam step with the kept gradient (bias-corrected; step count kept). */
fornos_status fornos_denoiser_adam(fornos_denoiser *d, float lr, float beta1, float beta2, float eps);
void fornos_denoiser_destroy(fornos_denoiser *d);

/* ---- the denoiser on a baked map ----------------------------------------- */
/* Network input image [height][width][FORNOS_DN_IN] (device memory, the map's
 * size) from the results of solvers on one mesh mapping whose texel set has
 * map indices: AO and bent normals (any sample count), position, normals
 * (object space) and height.  position, normals and height may all be null
 * (not some of them): the values those passes would give are then computed
 * from the mapping, which must have been run, within the gather, the same
 * bits without the three passes.  Positions are stored as
 * (p - center) * inv_extent, the mapping distance times dist_scale; texels
 * outside the set are zero.  image must be 16-byte aligned. */
fornos_status fornos_denoise_gather(const fornos_meshmap *mm, const fornos_solver *ao, const fornos_solver *bn,
                                    const fornos_solver *position, const fornos_solver *normals,
                                    const fornos_solver *height, const float *center, float inv_extent,
                                    float dist_scale, float *image);
/* Network output image [height][width][FORNOS_DN_OUT] back onto the texel
 * set: ao[n] clamped to [0, 1], bn[n][3] normalised (device memory). */
fornos_status fornos_denoise_scatter(const fornos_meshmap *mm, const float *image, float *ao, float *bn);
/* Consistency check per tile x tile block of the map: fail[ty][tx] = 1 when
 * the mean of (input AO - output AO) over the block's texels exceeds z
 * standard errors of a `samples`-sample AO estimate plus slack, i.e. the
 * network moved the block's mean further than the input's noise explains
 * (device memory; fail has ceil(h/tile) * ceil(w/tile) entries). */
fornos_status fornos_denoise_check_tiles(const float *input, const float *output, uint32_t width, uint32_t height,
                                         uint32_t tile, uint32_t samples, float z, float slack, uint32_t *fail);

/* ---- single operations (device pointers; bf16 = uint16_t bit patterns) --- */
/* One 3x3 convolution, as the network runs it: input channels [0, c1) from a
 * (half resolution and nearest-upsampled when up_a), [c1, c1 + c2) from b;
 * weights bf16 [cout][9][c1 + c2]; bias fp32 or null; relu; out fp32
 * [n][h][w][cout]. */
fornos_status fornos_dn_conv3x3(const uint16_t *a, uint32_t c1, int up_a, const uint16_t *b, uint32_t c2,
                                uint32_t n, uint32_t h, uint32_t w, const uint16_t *weights, const float *bias,
                                uint32_t cout, int relu, float *out);
/* Its weight and bias gradients for output gradient g (bf16 [n][h][w][cout]):
 * dw fp32 [cout][9][c1 + c2], db fp32 [cout] or null (both overwritten). */
fornos_status fornos_dn_conv3x3_wgrad(const uint16_t *a, uint32_t c1, int up_a, const uint16_t *b, uint32_t c2,
                                      uint32_t n, uint32_t h, uint32_t w, const uint16_t *g, uint32_t cout,
                                      float *dw, float *db);

#ifdef __cplusplus
}
#endif
#endif
