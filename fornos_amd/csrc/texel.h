// texel.h -- the per-texel passes (Shaders/heights.comp, positions.comp,
// normals.comp with tangentspace.comp fused).
#pragma once
#include "device.h"

namespace fornos {
namespace {

// ---------------------------------------------------------------------------
// Texel passes: heights.comp, positions.comp, normals.comp (+ tangentspace).
// ---------------------------------------------------------------------------
// heights.comp: the mapped hit distance, 0 where the texel missed
__device__ inline float texel_height(const float4* __restrict__ coords, uint32_t g)
{
    const float h = coords[g].x;
    return h != FLT_MAX ? h : 0.0f;
}

// positions.comp (NORMALS false) / normals.comp in object space (true): the
// hit triangle's vertex attribute interpolated at the texel's barycentrics,
// the normal normalised; zero where the texel missed
template <bool NORMALS>
__device__ inline V3 texel_attr(const float4* __restrict__ coords, const uint32_t* __restrict__ tidx,
                                const float4* __restrict__ soup, uint32_t g)
{
    const uint32_t t = tidx[g];
    if (t == kMiss) return mk(0, 0, 0);
    const float4 c = coords[g];
    const V3 r = xyz(soup[t]) * c.y + xyz(soup[t + 1]) * c.z + xyz(soup[t + 2]) * c.w;
    return NORMALS ? normalize(r) : r;
}

__global__ __launch_bounds__(kBlock) void k_height(const float4* __restrict__ coords, uint32_t begin,
                                                    uint32_t end, float* __restrict__ out)
{
    const uint32_t g = begin + blockIdx.x * kBlock + threadIdx.x;
    if (g >= end) return;
    out[g] = texel_height(coords, g);
}

template <bool NORMALS>
__global__ __launch_bounds__(kBlock) void k_attr(const float4* __restrict__ coords,
                                                  const uint32_t* __restrict__ tidx,
                                                  const float4* __restrict__ soup,
                                                  const float4* __restrict__ pixt, uint32_t begin,
                                                  uint32_t end, int tangent, float* __restrict__ out)
{
    const uint32_t g = begin + blockIdx.x * kBlock + threadIdx.x;
    if (g >= end) return;
    V3 r = texel_attr<NORMALS>(coords, tidx, soup, g);
    if (NORMALS && tangent && tidx[g] != kMiss) {
        const V3 n = xyz(pixt[3 * g]), tt = xyz(pixt[3 * g + 1]), b = xyz(pixt[3 * g + 2]);
        const V3 d0 = mk(n.z * b.y - n.y * b.z, n.x * b.z - n.z * b.x, n.y * b.x - n.x * b.y);
        const V3 d1 = mk(tt.z * n.y - tt.y * n.z, tt.x * n.z - n.x * tt.z, n.x * tt.y - tt.x * n.y);
        const V3 d2 = mk(tt.y * b.z - tt.z * b.y, tt.z * b.x - tt.x * b.z, tt.x * b.y - tt.y * b.x);
        r = normalize(mk(dot(r, d0), dot(r, d1), dot(r, d2)));
    }
    out[3 * g] = r.x;
    out[3 * g + 1] = r.y;
    out[3 * g + 2] = r.z;
}

}  // namespace
}  // namespace fornos
