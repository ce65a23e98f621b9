// mapping.h -- the mesh-mapping kernel (Shaders/meshmapping*.comp).
#pragma once
#include "device.h"

namespace fornos {
namespace {

// ---------------------------------------------------------------------------
// Mesh mapping: Shaders/meshmapping.comp (MODE 0: |nd| > 0) and
// meshmapping_nobackfaces.comp (MODE 1: forward nd > 0, MODE 2: backward
// nd < 0).  t bound: t < tmax for every mode (the non-culling shader tests
// t >= 0 and keeps strict-< minima, so skipping t >= current best is exact).
// ---------------------------------------------------------------------------
template <int MODE>
__device__ inline bool map_tri(V3 o, V3 d, uint32_t k, const DevScene& sc, float tmax, float& t_out,
                               V3& bc)
{
    const float4 q0 = sc.tri[k];
    const V3 n = xyz(q0);
    const float nd = dot(d, n);
    const bool side = MODE == 0 ? (fabsf(nd) > 0) : (MODE == 1 ? (nd > 0) : (nd < 0));
    if (!side) return false;
    const float pn = dot(o, n);
    const float t = (q0.w - pn) / nd;
    if (!(t >= 0 && t < tmax)) return false;
    const V3 pf = o + d * t;
    const float4 a4 = sc.spos[3 * k], b4 = sc.spos[3 * k + 1], c4 = sc.spos[3 * k + 2];
    // barycentric() in fp64, Shaders/meshmapping.comp:50-64
    const double ax = a4.x, ay = a4.y, az = a4.z;
    const double v0x = (double)b4.x - ax, v0y = (double)b4.y - ay, v0z = (double)b4.z - az;
    const double v1x = (double)c4.x - ax, v1y = (double)c4.y - ay, v1z = (double)c4.z - az;
    const double v2x = (double)pf.x - ax, v2y = (double)pf.y - ay, v2z = (double)pf.z - az;
    const double d00 = v0x * v0x + v0y * v0y + v0z * v0z;
    const double d01 = v0x * v1x + v0y * v1y + v0z * v1z;
    const double d11 = v1x * v1x + v1y * v1y + v1z * v1z;
    const double d20 = v2x * v0x + v2y * v0y + v2z * v0z;
    const double d21 = v2x * v1x + v2y * v1y + v2z * v1z;
    const double denom = d00 * d11 - d01 * d01;
    const double y = (d11 * d20 - d01 * d21) / denom;
    const double z = (d00 * d21 - d01 * d20) / denom;
    const float bx = (float)(1.0 - y - z), by = (float)y, bz = (float)z;
    const float lo = -1e-5f, hi = 1.0f;
    if (bx >= lo && by >= lo && by <= hi && bz >= lo && bz <= hi) {
        t_out = t;
        bc = mk(bx, by, bz);
        return true;
    }
    return false;
}

// One pass of raycastBVH / raycastBackBVH in the reference's stackless DFS
// order (Shaders/meshmapping_nobackfaces.comp:102-121): visit a node iff its
// slab entry is < the running best, else jump past its subtree.  The node
// test is node_visit: the reference's RayAABB(...) < cur decided on the
// reciprocal-multiply slab distances wherever they leave no doubt, with the
// reference's divisions for the rest.
template <int MODE>
__device__ inline void map_pass(const Ray& r, const DevScene& sc, float& cur, uint32_t& idx, V3& bc)
{
    const V3 o = r.o, d = r.d;
    uint32_t i = 0;
    while (i < sc.n_nodes) {
        const float4 lo = sc.nodes[2 * i], hi = sc.nodes[2 * i + 1];
        const uint32_t w = __float_as_uint(hi.w);
        if (node_visit(r, lo, hi, cur)) {
            if (w & kLeafFlag) {
                const uint32_t first = __float_as_uint(lo.w), end = first + (w & ~kLeafFlag);
                for (uint32_t k = first; k < end; ++k) {
                    float t;
                    V3 b;
                    if (map_tri<MODE>(o, d, k, sc, cur, t, b)) { cur = t; idx = 3 * k; bc = b; }
                }
            }
            ++i;
        } else {
            i = (w & kLeafFlag) ? i + 1 : w;
        }
    }
}

// Eight waves per SIMD (57 VGPRs, no spill; the compiler's own choice of 66
// leaves room for 7): more texels in flight hide the dependent node loads.
template <bool CULL>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_meshmap(const float4* __restrict__ pix, uint32_t begin,
                                                     uint32_t end, DevScene sc,
                                                     float4* __restrict__ coords,
                                                     uint32_t* __restrict__ tidx)
{
    const uint32_t g = begin + blockIdx.x * kBlock + threadIdx.x;
    if (g >= end) return;
    const V3 p = xyz(pix[2 * g]), d = xyz(pix[2 * g + 1]);
    uint32_t idx = kMiss;
    V3 bc = mk(0, 0, 0);
    float t = FLT_MAX;
    Ray fw, bw;
    fw.o = bw.o = p;
    fw.d = d;
    bw.d = -d;
    fw.inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    bw.inv = -fw.inv;  // 1/(-x) = -(1/x) exactly
    fw.oi = bw.oi = mk(0, 0, 0);  // (used by the culling test only)
    fw.slow = bw.slow = !(fabsf(d.x) >= kSlowDir && fabsf(d.y) >= kSlowDir && fabsf(d.z) >= kSlowDir);
    if (CULL) {
        // meshmapping_nobackfaces.comp:182-199: one running t shared by both passes
        map_pass<1>(fw, sc, t, idx, bc);
        map_pass<2>(bw, sc, t, idx, bc);
    } else {
        // meshmapping.comp:139-162: the backward pass starts from the forward t
        map_pass<0>(fw, sc, t, idx, bc);
        map_pass<0>(bw, sc, t, idx, bc);
    }
    coords[g] = make_float4(t, bc.x, bc.y, bc.z);
    tidx[g] = idx;
}

}  // namespace
}  // namespace fornos
