// raster.h -- the texel set on the GPU: MapUV raster (Src/compute.cpp:75-193)
// and CompressedMapUV (Src/compute.cpp:372-416).
//
// The reference rasterises triangles one after another and a later write
// overwrites an earlier one.  Here every (triangle, texel) pair that passes
// the barycentric test bids with atomicMax on the texel; the bid orders the
// writes the way the sequential loops do (triangle index first, then loop
// position: a write from the bound column x == W lands on the next row's
// x == 0 and precedes that row's own write).  A second pass recomputes the
// winner's barycentrics with the same arithmetic and writes the attributes,
// so the map equals the sequential one bit for bit.  Compaction keeps the
// texels with dot(d,d) > 0.5 in row-major order.
#pragma once
#include "device.h"

namespace fornos {
namespace {

// Per low-poly triangle: 3 corners x {p.xyz, uv.xy, n.xyz, dir.xyz, t.xyz, b.xyz}.
constexpr int kCornerFloats = 17;
constexpr int kTriFloats = 3 * kCornerFloats;

struct RasterArgs {
    const float* __restrict__ tris;   // kTriFloats per triangle
    uint32_t ntri, W, H;
    uint32_t* __restrict__ owner;     // W*H bids, 0 = none
    float* __restrict__ map;          // 5 x W*H*3 floats: pos, dir, nrm, tan, bit
};

struct RasterSetup {
    V2 u[3];
    uint32_t xMin, yMin, xMax, yMax;
};

__device__ inline uint32_t round_u32(float x) { return (uint32_t)(int64_t)roundf(x); }

__device__ inline V3 bary2(V2 p, V2 p0, V2 p1, V2 p2)  // Src/math.h:250-258
{
    const V2 v0 = {p1.x - p0.x, p1.y - p0.y}, v1 = {p2.x - p0.x, p2.y - p0.y}, v2 = {p.x - p0.x, p.y - p0.y};
    const float s = 1.0f / (v0.x * v1.y - v1.x * v0.y);
    const float i = (v2.x * v1.y - v1.x * v2.y) * s;
    const float j = (v0.x * v2.y - v2.x * v0.y) * s;
    return mk(1.0f - i - j, i, j);
}

__device__ inline RasterSetup raster_setup(const float* c, uint32_t W, uint32_t H)
{
    RasterSetup R;
    const V2 scale = {(float)W, (float)H};
    const V2 pixsize = {1.0f / scale.x, 1.0f / scale.y};
    const V2 halfpix = {pixsize.x * 0.5f, pixsize.y * 0.5f};
    V2 s[3];
    for (int k = 0; k < 3; ++k) {
        R.u[k] = {c[k * kCornerFloats + 3], c[k * kCornerFloats + 4]};
        s[k] = {(R.u[k].x - halfpix.x) * scale.x, (R.u[k].y - halfpix.y) * scale.y};
    }
    R.xMin = min(round_u32(fminf(s[0].x, fminf(s[1].x, s[2].x))), W);
    R.yMin = min(round_u32(fminf(s[0].y, fminf(s[1].y, s[2].y))), H);
    R.xMax = min(round_u32(fmaxf(s[0].x, fmaxf(s[1].x, s[2].x))), W);
    R.yMax = min(round_u32(fmaxf(s[0].y, fmaxf(s[1].y, s[2].y))), H);
    return R;
}

__device__ inline bool raster_inside(const RasterSetup& R, uint32_t x, uint32_t y, uint32_t W, uint32_t H, V3& b)
{
    const V2 pixsize = {1.0f / (float)W, 1.0f / (float)H};
    const V2 halfpix = {pixsize.x * 0.5f, pixsize.y * 0.5f};
    const V2 q = {(float)x * pixsize.x + halfpix.x, (float)y * pixsize.y + halfpix.y};
    b = bary2(q, R.u[0], R.u[1], R.u[2]);
    return b.x >= -0.001f && b.x <= 1 && b.y >= -0.001f && b.y <= 1 && b.z >= -0.001f && b.z <= 1;
}

// One block per triangle, texels of its bounding box spread over the block.
__global__ __launch_bounds__(kBlock) void k_raster_bid(RasterArgs A)
{
    const uint32_t t = blockIdx.x;
    if (t >= A.ntri) return;
    const RasterSetup R = raster_setup(A.tris + (size_t)t * kTriFloats, A.W, A.H);
    if (R.xMax < R.xMin || R.yMax < R.yMin) return;
    const uint32_t bw = R.xMax - R.xMin + 1, bh = R.yMax - R.yMin + 1;
    const uint64_t npx = (uint64_t)A.W * A.H;
    for (uint64_t k = threadIdx.x; k < (uint64_t)bw * bh; k += kBlock) {
        const uint32_t y = R.yMin + (uint32_t)(k / bw), x = R.xMin + (uint32_t)(k % bw);
        V3 b;
        if (!raster_inside(R, x, y, A.W, A.H, b)) continue;
        const uint64_t i = (uint64_t)y * A.W + x;
        if (i >= npx) continue;  // the reference writes past the end here (y == H)
        atomicMax(&A.owner[i], 2u * t + (x < A.W ? 2u : 1u));
    }
}

__device__ inline V3 blend3(const float* c, int off, V3 b)
{
    const V3 a0 = mk(c[off], c[off + 1], c[off + 2]);
    const V3 a1 = mk(c[kCornerFloats + off], c[kCornerFloats + off + 1], c[kCornerFloats + off + 2]);
    const V3 a2 = mk(c[2 * kCornerFloats + off], c[2 * kCornerFloats + off + 1], c[2 * kCornerFloats + off + 2]);
    return a0 * b.x + a1 * b.y + a2 * b.z;
}

// Writes the winning triangle's attributes (Src/compute.cpp:152-157).
__global__ __launch_bounds__(kBlock) void k_raster_write(RasterArgs A)
{
    const uint64_t npx = (uint64_t)A.W * A.H;
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= npx) return;
    const uint32_t key = A.owner[i];
    if (key == 0) return;
    const uint32_t t = (key - 1) / 2;
    const bool bound_column = ((key - 1) & 1u) == 0;  // written from x == W of the row above
    const uint32_t x = bound_column ? A.W : (uint32_t)(i % A.W);
    const uint32_t y = bound_column ? (uint32_t)(i / A.W) - 1 : (uint32_t)(i / A.W);
    const float* c = A.tris + (size_t)t * kTriFloats;
    const RasterSetup R = raster_setup(c, A.W, A.H);
    V3 b;
    raster_inside(R, x, y, A.W, A.H, b);
    float* m = A.map;
    st3(m + 0 * 3 * npx, i, blend3(c, 0, b));
    st3(m + 1 * 3 * npx, i, normalize(blend3(c, 8, b)));
    st3(m + 2 * 3 * npx, i, normalize(blend3(c, 5, b)));
    st3(m + 3 * 3 * npx, i, normalize(blend3(c, 11, b)));
    st3(m + 4 * 3 * npx, i, normalize(blend3(c, 14, b)));
}

// Texels with direction data (Src/compute.cpp:379-386): flag per texel.
__global__ __launch_bounds__(kBlock) void k_texel_flags(const float* __restrict__ dir, uint64_t npx,
                                                         uint32_t* __restrict__ flags)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= npx) return;
    const V3 n = ld3(dir, i);
    flags[i] = dot(n, n) > 0.5f ? 1u : 0u;
}

// Scatter of the compressed texel set into the mapping inputs (pix: p, d;
// pixt: n, t, b as float4) plus the MapUV index of each texel.
__global__ __launch_bounds__(kBlock) void k_texel_compact(const float* __restrict__ map, uint64_t npx,
                                                           const uint32_t* __restrict__ flags,
                                                           const uint32_t* __restrict__ slot,
                                                           float4* __restrict__ pix, float4* __restrict__ pixt,
                                                           uint32_t* __restrict__ indices)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= npx || !flags[i]) return;
    const uint32_t k = slot[i];
    const float* pos = map;
    const float* dir = map + 3 * npx;
    const float* nrm = map + 6 * npx;
    const float* tan = map + 9 * npx;
    const float* bit = map + 12 * npx;
    pix[2 * (size_t)k] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.0f);
    pix[2 * (size_t)k + 1] = make_float4(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2], 0.0f);
    pixt[3 * (size_t)k] = make_float4(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2], 0.0f);
    pixt[3 * (size_t)k + 1] = make_float4(tan[3 * i], tan[3 * i + 1], tan[3 * i + 2], 0.0f);
    pixt[3 * (size_t)k + 2] = make_float4(bit[3 * i], bit[3 * i + 1], bit[3 * i + 2], 0.0f);
    indices[k] = (uint32_t)i;
}

}  // namespace
}  // namespace fornos
