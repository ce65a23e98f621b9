/*
 * fornos_export.h -- C ABI of the image export of a baked map (the
 * reference's Src/image.cpp), on the solver's device results.
 *
 * Orientation as the reference writes it: texel index idx = y * w + x of the
 * w x h map lands in image row h - y - 1.  RGB8 buffers are interleaved
 * (w * h * 3 bytes), float buffers are planes of w * h floats.  The solver's
 * mesh mapping must carry map indices (a texel set built with indices or on
 * the device).  Output pointers are host memory.
 *
 * Exports run on the scene's export stream after the solver's last queued
 * pass and return when their results are in host memory; other passes may
 * run meanwhile (a host thread exporting one map while the next is baked).
 * Exports of one scene take turns; a solver must not run again while it is
 * being exported.
 */
#ifndef FORNOS_EXPORT_H
#define FORNOS_EXPORT_H

#include "fornos_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* exportFloatImage, 8-bit branch (Src/image.cpp:185-240): value range over
 * the texels (NaN ignored), clamped to [filter_lo, filter_hi] when
 * filter_hi > filter_lo; scale and bias from it when normalize; values
 * outside the filter range become 0; t * scale + bias clamped to [0, 1] and
 * truncated to 0..255, written as gray RGB; dilate > 0: dilateRGB over the
 * texel set (Src/image.cpp:127-183) with that distance.  *minmax (2 floats,
 * may be null) = the range after the filter. */
fornos_status fornos_export_float_rgb8(const fornos_solver *s, float filter_lo, float filter_hi, int normalize,
                                       int dilate, uint8_t *rgb, float *minmax);
/* exportFloatImage, EXR branch (Src/image.cpp:241-277): the filter clamps;
 * t = normalize ? d * scale + bias : d; one plane of w * h floats. */
fornos_status fornos_export_float_f32(const fornos_solver *s, float filter_lo, float filter_hi, int normalize,
                                      float *image, float *minmax);
/* exportNormalImage, 8-bit branch (Src/image.cpp:352-386): n * 0.5 + 0.5
 * times 255 truncated per component; dilation over texels whose colour is
 * not (0, 0, 0) (the row of a texel is idx / w; the reference divides by h,
 * which differs for non-square maps only). */
fornos_status fornos_export_normal_rgb8(const fornos_solver *s, int dilate, uint8_t *rgb);
/* exportVectorImage (Src/image.cpp:280-342): planes B, G, R = z, y, x. */
fornos_status fornos_export_vector_f32(const fornos_solver *s, float *planes);
/* Width and height of the map a solver's texel set belongs to. */
fornos_status fornos_solver_map_size(const fornos_solver *s, uint32_t *width, uint32_t *height);

/* The reference's three export functions with their extension dispatch:
 * .png / .tga (8-bit) or .exr (half channels, no compression).  An unknown
 * extension is an error here (the reference returns silently). */
fornos_status fornos_export_float_image(const fornos_solver *s, const char *path, float filter_lo, float filter_hi,
                                        int normalize, int dilate, float *minmax);
fornos_status fornos_export_normal_image(const fornos_solver *s, const char *path, int dilate);
fornos_status fornos_export_vector_image(const fornos_solver *s, const char *path);

/* Writers (host memory). */
fornos_status fornos_write_png(const char *path, uint32_t width, uint32_t height, const uint8_t *rgb);
fornos_status fornos_write_tga(const char *path, uint32_t width, uint32_t height, const uint8_t *rgb);
/* channels half planes, names sorted (EXR requires it), increasing y. */
fornos_status fornos_write_exr(const char *path, uint32_t width, uint32_t height, uint32_t channels,
                               const char *const *names, const float *const *planes);
/* tinyexr's float -> half conversion (3rdParty/tinyexr/tinyexr.h:7044-7080). */
void fornos_float_to_half(const float *in, uint16_t *out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
