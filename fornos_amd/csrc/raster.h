// raster.h -- the texel set on the GPU: MapUV raster (Src/compute.cpp:75-193,
// and the Hybrid method's createMapUVEdge, :213-349) and CompressedMapUV
// (Src/compute.cpp:372-416).
//
// The reference rasterises triangles one after another and a later write
// overwrites an earlier one.  Here every (triangle, texel) pair that passes
// the barycentric test bids with a 64-bit atomicMax on the texel: the bid is
// (triangle + 1) << 32 | row, which orders the writes the way the
// sequential loops do (triangle index first; within a triangle two loop
// positions that land on one texel differ in row, and the later row wins,
// as when a write from column x >= W lands on a following row).  A second
// pass recomputes the winner's barycentrics with the same arithmetic and
// writes the attributes, so the map equals the sequential one bit for bit.
// Compaction keeps the texels with dot(d,d) > 0.5 in row-major order.
#pragma once
#include "device.h"
#include "mapuv.h"

namespace fornos {
namespace {

// Per low-poly triangle: 3 corners x {p.xyz, uv.xy, n.xyz, dir.xyz, t.xyz, b.xyz}.
constexpr int kCornerFloats = 17;
constexpr int kTriFloats = 3 * kCornerFloats;

struct RasterArgs {
    const float* __restrict__ tris;   // kTriFloats per triangle
    uint32_t ntri, W, H;
    float edge;                       // > 0: the Hybrid method
    unsigned long long* __restrict__ owner;  // W*H bids, 0 = none
    float* __restrict__ map;          // 5 x W*H*3 floats: pos, dir, nrm, tan, bit
};

__device__ inline void corner_uv(const float* c, V2* u)
{
    for (int k = 0; k < 3; ++k) u[k] = {c[k * kCornerFloats + 3], c[k * kCornerFloats + 4]};
}

// One block per triangle, texels of its bounding box spread over the block.
__global__ __launch_bounds__(kBlock) void k_raster_bid(RasterArgs A)
{
    const uint32_t t = blockIdx.x;
    if (t >= A.ntri) return;
    V2 u[3];
    corner_uv(A.tris + (size_t)t * kTriFloats, u);
    const TexelBox R = texel_box(u, A.W, A.H, A.edge > 0.0f);
    if (!R.ok || R.xMax < R.xMin || R.yMax < R.yMin) return;
    const uint32_t bw = R.xMax - R.xMin + 1, bh = R.yMax - R.yMin + 1;
    const uint64_t npx = (uint64_t)A.W * A.H;
    for (uint64_t k = threadIdx.x; k < (uint64_t)bw * bh; k += kBlock) {
        const uint32_t y = R.yMin + (uint32_t)(k / bw), x = R.xMin + (uint32_t)(k % bw);
        V3 b;
        if (!texel_inside(u, x, y, A.W, A.H, b)) continue;
        const uint64_t i = (uint64_t)y * A.W + x;
        if (i >= npx) continue;  // the reference writes past the end here (y == H)
        atomicMax(&A.owner[i], ((unsigned long long)(t + 1) << 32) | y);
    }
}

__device__ inline V3 blend3(const float* c, int off, V3 b)
{
    const V3 a0 = mk(c[off], c[off + 1], c[off + 2]);
    const V3 a1 = mk(c[kCornerFloats + off], c[kCornerFloats + off + 1], c[kCornerFloats + off + 2]);
    const V3 a2 = mk(c[2 * kCornerFloats + off], c[2 * kCornerFloats + off + 1], c[2 * kCornerFloats + off + 2]);
    return a0 * b.x + a1 * b.y + a2 * b.z;
}

// Writes the winning triangle's attributes (Src/compute.cpp:152-157; the
// Hybrid direction and tangent frame, :294-313).
__global__ __launch_bounds__(kBlock) void k_raster_write(RasterArgs A)
{
    const uint64_t npx = (uint64_t)A.W * A.H;
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= npx) return;
    const unsigned long long key = A.owner[i];
    if (key == 0) return;
    const uint32_t t = (uint32_t)(key >> 32) - 1u, y = (uint32_t)key;
    const uint32_t x = (uint32_t)(i - (uint64_t)y * A.W);
    const float* c = A.tris + (size_t)t * kTriFloats;
    V2 u[3];
    corner_uv(c, u);
    V3 b;
    texel_inside(u, x, y, A.W, A.H, b);
    float* m = A.map;
    const V3 p = blend3(c, 0, b), n = normalize(blend3(c, 5, b));
    V3 d = normalize(blend3(c, 8, b));
    if (A.edge > 0.0f) {
        const V3 p0 = mk(c[0], c[1], c[2]);
        const V3 p1 = mk(c[kCornerFloats], c[kCornerFloats + 1], c[kCornerFloats + 2]);
        const V3 p2 = mk(c[2 * kCornerFloats], c[2 * kCornerFloats + 1], c[2 * kCornerFloats + 2]);
        d = hybrid_direction(p0, p1, p2, p, n, d, A.edge);
    }
    st3(m + 0 * 3 * npx, i, p);
    st3(m + 1 * 3 * npx, i, d);
    st3(m + 2 * 3 * npx, i, n);
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
