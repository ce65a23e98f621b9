// export.h -- device side of the image export (Src/image.cpp): the value
// range, the 8-bit quantisation, the float image and the dilation of the
// baked map, all on the texel results where the solver left them.
//
// Layout as the reference writes it: pixel (x, y) of the map (texel index
// idx = y * w + x) lands in row h - y - 1 of the image; RGB8 images are
// interleaved, float images one plane per channel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace fornos {
namespace {

// std::max(t, 0.0f), std::min(t, 1.0f) as the reference evaluates them
__device__ inline float std_max0(float t) { return (t < 0.0f) ? 0.0f : t; }
__device__ inline float std_min1(float t) { return (1.0f < t) ? 1.0f : t; }

// (uint8_t) of a float on the reference's x86 targets: truncating 32-bit
// conversion (NaN and out of range give 0x80000000), low byte kept
__device__ inline uint8_t to_u8(float v)
{
    const int32_t i = (isnan(v) || v >= 2147483648.0f || v < -2147483648.0f) ? (int32_t)0x80000000 : (int32_t)v;
    return (uint8_t)(i & 0xff);
}

// order-preserving key of a float for integer atomics
__device__ inline uint32_t fkey(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float fkey_inv(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// getMinMax (Src/image.cpp:37-47): fminf / fmaxf over the texels, NaN ignored.
// keys[0] = min key (start 0xffffffff), keys[1] = max key (start 0).
__global__ void k_minmax(const float* v, uint32_t n, uint32_t* keys)
{
    uint32_t lo = 0xffffffffu, hi = 0u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float f = v[i];
        if (isnan(f)) continue;
        const uint32_t k = fkey(f);
        lo = min(lo, k);
        hi = max(hi, k);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, off));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&keys[0], lo);
        atomicMax(&keys[1], hi);
    }
}

struct QuantArgs {
    const float* v;          // n values (1 per texel) or 3n (vectors)
    const uint32_t* indices; // map pixel of each texel
    uint32_t n, w, h;
    float lo, hi, scale, bias;
    int filter, normalize;
};

// exportFloatImage, PNG / TGA branch (Src/image.cpp:212-233): values outside
// the filter range become 0, then t = d * scale + bias, clamped to [0, 1],
// times 255, truncated; gray RGB.
__global__ void k_quant_gray(QuantArgs A, uint8_t* rgb)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const float d = A.v[i];
    const float df = (A.filter && (d < A.lo || d > A.hi)) ? 0.0f : d;
    const float t = df * A.scale + A.bias;
    const uint8_t c = to_u8(std_min1(std_max0(t)) * 255.0f);
    const uint32_t idx = A.indices[i], x = idx % A.w, y = idx / A.w;
    const size_t p = ((size_t)(A.h - y - 1) * A.w + x) * 3;
    rgb[p + 0] = c;
    rgb[p + 1] = c;
    rgb[p + 2] = c;
}

// exportFloatImage, EXR branch (Src/image.cpp:249-260): the filter clamps,
// normalisation is optional.
__global__ void k_float_plane(QuantArgs A, float* img)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const float d = A.v[i];
    const float df = A.filter ? fminf(fmaxf(d, A.lo), A.hi) : d;
    const float t = A.normalize ? df * A.scale + A.bias : df;
    const uint32_t idx = A.indices[i], x = idx % A.w, y = idx / A.w;
    img[(size_t)(A.h - y - 1) * A.w + x] = t;
}

// exportNormalImage, PNG / TGA branch (Src/image.cpp:372-381): n * 0.5 + 0.5,
// times 255, truncated, per component.
__global__ void k_quant_normal(QuantArgs A, uint8_t* rgb)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const uint32_t idx = A.indices[i], x = idx % A.w, y = idx / A.w;
    const size_t p = ((size_t)(A.h - y - 1) * A.w + x) * 3;
    for (int k = 0; k < 3; ++k) {
        const float c = A.v[3 * (size_t)i + k] * 0.5f + 0.5f;
        rgb[p + k] = to_u8(c * 255.0f);
    }
}

// exportVectorImage (Src/image.cpp:299-306): planes B, G, R = z, y, x.
__global__ void k_vector_planes(QuantArgs A, float* planes)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const uint32_t idx = A.indices[i], x = idx % A.w, y = idx / A.w;
    const size_t p = (size_t)(A.h - y - 1) * A.w + x, plane = (size_t)A.w * A.h;
    planes[p] = A.v[3 * (size_t)i + 2];
    planes[plane + p] = A.v[3 * (size_t)i + 1];
    planes[2 * plane + p] = A.v[3 * (size_t)i + 0];
}

// Valid-pixel tables (map order, not flipped).  createValidPixelsTable
// (Src/image.cpp:92-100): every texel; createValidPixelsTableRGB (:102-123):
// texels whose RGB is not (0, 0, 0) (the row of a texel is idx / w).
__global__ void k_valid_texels(const uint32_t* indices, uint32_t n, uint8_t* valid)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) valid[indices[i]] = 1;
}

__global__ void k_valid_rgb(const uint32_t* indices, uint32_t n, uint32_t w, uint32_t h, const uint8_t* rgb,
                            uint8_t* valid)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t idx = indices[i], x = idx % w, y = idx / w;
    const size_t p = ((size_t)(h - y - 1) * w + x) * 3;
    valid[idx] = (rgb[p] != 0 || rgb[p + 1] != 0 || rgb[p + 2] != 0) ? 1 : 0;
}

// dilateRGB (Src/image.cpp:127-183) as it runs on one thread: the pixels in
// the map's row-major order (y, then x), each invalid one filled in place.
// Pass 1, one thread per pixel: an invalid pixel's centre, the first
// pos + offset * dist (dist = 1 .. max_dist - 1, the eight offsets in the
// reference's order) strictly inside the border whose eight neighbours are all
// valid, as a map index (kNoCentre: none).
constexpr uint32_t kNoCentre = 0xffffffffu;

__global__ void k_dilate_centre(const uint8_t* valid, uint32_t* centre, uint32_t w, uint32_t h, uint32_t max_dist)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    const size_t p = (size_t)y * w + x;
    centre[p] = kNoCentre;
    if (valid[p]) return;
    const int ox[8] = {1, -1, 0, 0, 1, 1, -1, -1}, oy[8] = {0, 0, 1, -1, 1, -1, 1, -1};
    const int W = (int)w, H = (int)h;
    for (int dist = 1; dist < (int)max_dist; ++dist) {
        for (int k = 0; k < 8; ++k) {
            const int cx = (int)x + ox[k] * dist, cy = (int)y + oy[k] * dist;
            if (!(cx > 0 && cy > 0 && cx < W - 1 && cy < H - 1)) continue;
            bool ok = true;
            for (int j = 0; j < 8 && ok; ++j) ok = valid[(size_t)(cy + oy[j]) * w + (cx + ox[j])] != 0;
            if (!ok) continue;
            centre[p] = (uint32_t)cy * w + (uint32_t)cx;
            return;
        }
    }
}

// Pass 2: the colour an invalid pixel ends with.  It copies its centre c as
// the in-place loop finds it: c's final colour when c is an invalid pixel the
// loop has already passed (only an isolated hole can be a centre, since a
// centre's eight neighbours are valid), else c's colour before dilation.  The
// chain moves to ever earlier pixels, so it ends.  dst starts as a copy of src.
__global__ void k_dilate_fill(const uint8_t* src, uint8_t* dst, const uint8_t* valid, const uint32_t* centre,
                              uint32_t w, uint32_t h)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    const uint32_t p = y * w + x;
    uint32_t c = centre[p];
    if (c == kNoCentre) return;
    uint32_t q = p;
    while (!valid[c] && c < q) {
        const uint32_t n = centre[c];
        if (n == kNoCentre) break;  // c was never filled: its colour before dilation
        q = c;
        c = n;
    }
    const uint32_t cy = c / w, cx = c - cy * w;
    const size_t s = ((size_t)(h - cy - 1) * w + cx) * 3, d = ((size_t)(h - y - 1) * w + x) * 3;
    dst[d + 0] = src[s + 0];
    dst[d + 1] = src[s + 1];
    dst[d + 2] = src[s + 2];
}

}  // namespace
}  // namespace fornos
