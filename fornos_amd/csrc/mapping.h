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

// The same triangle with a watertight containment test in place of the
// reference's plane distance + fp64 barycentrics with their -1e-5 tolerance:
// the ray/triangle method of Woop, Benthin and Wald ("Watertight Ray/Triangle
// Intersection", JCGT 2(1), 2013).  The vertices, relative to the origin, are
// sheared into the ray's frame (z along the dominant direction component),
// the three 2-D edge functions U, V, W are evaluated there (recomputed in
// double when one is exactly zero), and the ray hits iff they share a sign:
// a ray through an edge or vertex shared by two triangles hits at least one
// of them.  t = (U Az + V Bz + W Cz) / (U + V + W); the barycentrics are
// (U, V, W) / (U + V + W).  The side rule (culling) is the reference's n.d
// test.  No reference counterpart: the production option of the mapping.
template <int MODE>
__device__ inline bool map_tri_wt(V3 o, V3 d, uint32_t k, const DevScene& sc, float tmax, float& t_out, V3& bc)
{
    const float4 q0 = sc.tri[k];
    const float nd = dot(d, xyz(q0));
    const bool side = MODE == 0 ? (fabsf(nd) > 0) : (MODE == 1 ? (nd > 0) : (nd < 0));
    if (!side) return false;
    const float dd[3] = {d.x, d.y, d.z};
    const float ad[3] = {fabsf(d.x), fabsf(d.y), fabsf(d.z)};
    const int kz = ad[0] > ad[1] ? (ad[0] > ad[2] ? 0 : 2) : (ad[1] > ad[2] ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
    if (dd[kz] < 0.0f) {  // keep the winding
        const int t = kx;
        kx = ky;
        ky = t;
    }
    const float sx = dd[kx] / dd[kz], sy = dd[ky] / dd[kz], sz = 1.0f / dd[kz];
    const float4 a4 = sc.spos[3 * k], b4 = sc.spos[3 * k + 1], c4 = sc.spos[3 * k + 2];
    const float A[3] = {a4.x - o.x, a4.y - o.y, a4.z - o.z};
    const float B[3] = {b4.x - o.x, b4.y - o.y, b4.z - o.z};
    const float C[3] = {c4.x - o.x, c4.y - o.y, c4.z - o.z};
    const float ax = A[kx] - sx * A[kz], ay = A[ky] - sy * A[kz];
    const float bx = B[kx] - sx * B[kz], by = B[ky] - sy * B[kz];
    const float cx = C[kx] - sx * C[kz], cy = C[ky] - sy * C[kz];
    float U = cx * by - cy * bx, V = ax * cy - ay * cx, W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)cx * by - (double)cy * bx);
        V = (float)((double)ax * cy - (double)ay * cx);
        W = (float)((double)bx * ay - (double)by * ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T = U * (sz * A[kz]) + V * (sz * B[kz]) + W * (sz * C[kz]);
    const float t = T / det;
    if (!(t >= 0 && t < tmax)) return false;
    t_out = t;
    bc = mk(U / det, V / det, W / det);
    return true;
}

// One pass of raycastBVH / raycastBackBVH in the reference's stackless DFS
// order (Shaders/meshmapping_nobackfaces.comp:102-121): visit a node iff its
// slab entry is < the running best, else jump past its subtree.  The node
// test is node_visit: the reference's RayAABB(...) < cur decided on the
// reciprocal-multiply slab distances wherever they leave no doubt, with the
// reference's divisions for the rest.  WT: the conservative node_visit_wt and
// the watertight triangle test.
template <int MODE, bool WT>
__device__ inline void map_pass(const Ray& r, const DevScene& sc, float& cur, uint32_t& idx, V3& bc)
{
    const V3 o = r.o, d = r.d;
    uint32_t i = 0;
    while (i < sc.n_nodes) {
        const float4 lo = sc.nodes[2 * i], hi = sc.nodes[2 * i + 1];
        const uint32_t w = __float_as_uint(hi.w);
        if (WT ? node_visit_wt(r, lo, hi, cur) : node_visit(r, lo, hi, cur)) {
            if (w & kLeafFlag) {
                const uint32_t first = __float_as_uint(lo.w), end = first + (w & ~kLeafFlag);
                for (uint32_t k = first; k < end; ++k) {
                    float t;
                    V3 b;
                    const bool hit = WT ? map_tri_wt<MODE>(o, d, k, sc, cur, t, b) : map_tri<MODE>(o, d, k, sc, cur, t, b);
                    if (hit) { cur = t; idx = 3 * k; bc = b; }
                }
            }
            ++i;
        } else {
            i = (w & kLeafFlag) ? i + 1 : w;
        }
    }
}

// One pass on the 4-wide culling tree with the reference pass's result.  The
// reference pass (map_pass) keeps, among the triangles of the leaves it
// reaches, the first in its DFS visit order with the smallest t; it reaches a
// leaf iff every node of the leaf's chain has slab entry < its running best
// when tested, and that running best starts at `cur0` (FLT_MAX, or the
// forward pass's result) and never drops below the pass's result t*.  The
// soup holds the triangles in the DFS order of their leaves, so the first
// visited is the one with the smallest soup index.  Here:
//  - a box is skipped only when the ray enters it no earlier than the best hit
//    so far: every hit the mapping's tolerance accepts lies strictly inside
//    its culling box (the boxes' margin), so a skipped box holds no hit that
//    beats or ties the best;
//  - a hit that beats the best, or ties it with a smaller index (within this
//    pass), is taken when its leaf's chain passes with bound t.  Taken hits
//    never beat the reference's: a taken t < t* would have a chain passing
//    with t < t* <= the reference's running best, so the reference would
//    reach the leaf and find it.  And the reference's own winner is taken
//    whenever its chain passes with its t;
//  - when the chain passes only with a larger bound (up to cur0), whether the
//    reference reached the leaf depends on its visit order: the pass is redone
//    on the reference tree, as when the stack would overflow.
// Returns 0 when the pass is decided, kMapRedoChain + 1 or kMapRedoStack + 1
// to redo it (cur, idx and bc are then partial; the caller restores them).
enum { kMapRedoChain = 0, kMapRedoStack = 1 };  // counters of passes redone on the reference tree
template <int MODE>
__device__ inline int map_wide_pass(const Ray& r, const DevScene& sc, uint32_t* stk, uint32_t depth, float cur0,
                                    float& cur, uint32_t& idx, V3& bc)
{
    const uint32_t tid = threadIdx.x;
    bool mine = false;  // the best hit is this pass's own
    uint32_t sp = 0, node = 0;
    for (;;) {
        if (node == kNoChild) {
            if (sp == 0) return 0;
            --sp;
            node = stk[sp * kBlock + tid];
        }
        if (node & kLeafFlag) {
            const uint4 lf = sc.wleaf[node & ~kLeafFlag];
            node = kNoChild;
            for (uint32_t k = lf.x; k < lf.x + lf.y; ++k) {
                float t;
                V3 b;
                // ties: a smaller index of this pass wins, so let t == cur
                // through (cur >= 0 here: the next float up is one ulp above)
                const float tmax = mine ? __uint_as_float(__float_as_uint(cur) + 1u) : cur;
                if (!map_tri<MODE>(r.o, r.d, k, sc, tmax, t, b)) continue;
                if (t == cur && !(3 * k < idx)) continue;
                if (chain_ok(r, sc, lf.z, t)) {
                    cur = t;
                    idx = 3 * k;
                    bc = b;
                    mine = true;
                } else if (chain_ok(r, sc, lf.z, cur0)) {
                    return kMapRedoChain + 1;
                }
            }
            continue;
        }
        const float4* nd = sc.wide + 8 * (size_t)node;
        const float4 lx = nd[0], ly = nd[1], lz = nd[2], hx = nd[3], hy = nd[4], hz = nd[5];
        const uint4 ch = *(const uint4*)(nd + 6);
        const float4 lo[4] = {make_float4(lx.x, ly.x, lz.x, 0), make_float4(lx.y, ly.y, lz.y, 0),
                              make_float4(lx.z, ly.z, lz.z, 0), make_float4(lx.w, ly.w, lz.w, 0)};
        const float4 hi[4] = {make_float4(hx.x, hy.x, hz.x, 0), make_float4(hx.y, hy.y, hz.y, 0),
                              make_float4(hx.z, hy.z, hz.z, 0), make_float4(hx.w, hy.w, hz.w, 0)};
        const uint32_t chs[4] = {ch.x, ch.y, ch.z, ch.w};
        float e[4];
        uint32_t c[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float al = 0;
            const bool pass = (chs[k] != kNoChild) & box_maybe(r, lo[k], hi[k], cur, al);
            e[k] = pass ? al : INFINITY;
            c[k] = pass ? chs[k] : kNoChild;
        }
#define FORNOS_CSWAP(i, j)                                            \
    if (e[j] < e[i]) {                                                \
        const float te = e[i]; e[i] = e[j]; e[j] = te;                \
        const uint32_t tc = c[i]; c[i] = c[j]; c[j] = tc;             \
    }
        FORNOS_CSWAP(0, 1) FORNOS_CSWAP(2, 3) FORNOS_CSWAP(0, 2) FORNOS_CSWAP(1, 3) FORNOS_CSWAP(1, 2)
#undef FORNOS_CSWAP
        node = c[0];
        const uint32_t n = (uint32_t)(c[1] != kNoChild) + (uint32_t)(c[2] != kNoChild) + (uint32_t)(c[3] != kNoChild);
        if (sp + n > depth) return kMapRedoStack + 1;
#pragma unroll
        for (uint32_t k = 1; k <= 3; ++k)
            if (k <= n) stk[(sp + n - k) * kBlock + tid] = c[k];
        sp += n;
    }
}

constexpr int kMapStack = 24;  // 4-wide mapping stack entries per lane (LDS); `depth` may use fewer

// The reference predicate on the 4-wide culling tree (map_wide_pass), each
// pass redone on the reference tree when it cannot be decided there, and rays
// with a direction component below 2^-40 traced on the reference tree from
// the start.  Results equal k_meshmap<CULL, false>'s bit for bit.
template <bool CULL>
__global__ __launch_bounds__(kBlock) void k_meshmap_wide(const float4* __restrict__ pix, uint32_t begin, uint32_t end,
                                                        DevScene sc, float4* __restrict__ coords,
                                                        uint32_t* __restrict__ tidx, uint32_t* __restrict__ redo,
                                                        uint32_t depth)
{
    __shared__ uint32_t stk[kMapStack * kBlock];
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
    fw.oi = mk(p.x * fw.inv.x, p.y * fw.inv.y, p.z * fw.inv.z);
    bw.oi = -fw.oi;    // RN(o * (-inv)) = -RN(o * inv)
    fw.slow = bw.slow = !(fabsf(d.x) >= kSlowDir && fabsf(d.y) >= kSlowDir && fabsf(d.z) >= kSlowDir);
    constexpr int M1 = CULL ? 1 : 0, M2 = CULL ? 2 : 0;
    if (fw.slow) {
        map_pass<M1, false>(fw, sc, t, idx, bc);
        map_pass<M2, false>(bw, sc, t, idx, bc);
    } else {
        // meshmapping_nobackfaces.comp:182-199 / meshmapping.comp:139-162:
        // the backward pass starts from the forward pass's result
        float t0 = t;
        uint32_t i0 = idx;
        V3 b0 = bc;
        if (const int why = map_wide_pass<M1>(fw, sc, stk, depth, t0, t, idx, bc)) {
            t = t0; idx = i0; bc = b0;
            map_pass<M1, false>(fw, sc, t, idx, bc);
            atomicAdd(&redo[why - 1], 1u);
        }
        t0 = t; i0 = idx; b0 = bc;
        if (const int why = map_wide_pass<M2>(bw, sc, stk, depth, t0, t, idx, bc)) {
            t = t0; idx = i0; bc = b0;
            map_pass<M2, false>(bw, sc, t, idx, bc);
            atomicAdd(&redo[why - 1], 1u);
        }
    }
    coords[g] = make_float4(t, bc.x, bc.y, bc.z);
    tidx[g] = idx;
}

// Eight waves per SIMD (57 VGPRs, no spill; the compiler's own choice of 66
// leaves room for 7): more texels in flight hide the dependent node loads.
// WT: the watertight node and triangle tests (node_visit_wt, map_tri_wt)
// instead of the reference's.
template <bool CULL, bool WT>
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
        map_pass<1, WT>(fw, sc, t, idx, bc);
        map_pass<2, WT>(bw, sc, t, idx, bc);
    } else {
        // meshmapping.comp:139-162: the backward pass starts from the forward t
        map_pass<0, WT>(fw, sc, t, idx, bc);
        map_pass<0, WT>(bw, sc, t, idx, bc);
    }
    coords[g] = make_float4(t, bc.x, bc.y, bc.z);
    tidx[g] = idx;
}

}  // namespace
}  // namespace fornos
