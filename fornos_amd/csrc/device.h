// device.h -- device-side data layout and the arithmetic shared by every
// kernel: the reference's slab test and triangle predicates, the ray basis
// and sample directions.  Included once, by kernels.hip.
//
// HBM layout of a scene (all soup-ordered, i.e. in the DFS leaf order of the
// reference flatten, Src/meshmapping.cpp:99-116):
//   nodes  2 x float4 per node: (min.xyz, first soup triangle),
//                               (max.xyz, jump | kLeafFlag+count)
//   tri    (n.xyz, dot(a,n)) per triangle, then 3 x float4 per triangle:
//          (a.xyz, d00), (v0.xyz, d01), (v1.xyz, d11), with
//          n = normalize(cross(b-a, c-a)), v0 = b-a, v1 = c-a -- the
//          per-triangle terms of the predicates, computed once on the host in
//          the shader's own operation order
//   spos   3 x float4 per triangle (reference Vector4 soup positions)
//   snrm   3 x float4 per triangle (reference Vector4 soup normals)
//   chain  reachability links of the reference tree (the nearest ancestor
//          whose test a node's does not imply); cnodes / wide / wleaf: the
//          culling trees
// Texel data: pix 2 x float4 per texel (p, d), pixt 3 x float4 (n, t, b),
// coords float4 (t, b0, b1, b2), tidx uint (soup vertex index).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>

#include "host.h"

#pragma clang fp contract(off)

namespace fornos {
namespace {

constexpr uint32_t kMiss = 0xFFFFFFFFu;
constexpr uint32_t kWorkPerStep = 1024 * 128;  // k_workPerFrame, Src/solver_ao.cpp:33
constexpr uint32_t kGroup = 64;                // k_groupSize, Src/solver_ao.cpp:32
constexpr int kBlock = 256;

struct DevScene {
    const float4* __restrict__ nodes;    // reference tree
    const float4* __restrict__ tri;      // (n, a.n) per triangle
    const float4* __restrict__ trest;    // (a, d00), (v0, d01), (v1, d11) per triangle
    const float4* __restrict__ spos;
    const float4* __restrict__ snrm;
    const uint32_t* __restrict__ chain;  // reference tree reachability links
    const float4* __restrict__ cnodes;   // culling tree over the reference leaves
    const uint32_t* __restrict__ cref;   // reference leaf of each culling leaf
    const float4* __restrict__ wide;     // the culling tree collapsed to 4 children, 8 float4 per node
    const uint4* __restrict__ wleaf;     // its leaves: first soup triangle, count, reference leaf, group node + 1 or 0
    const float4* __restrict__ wide_ao;  // AO's 4-wide tree (leaves may be triangle groups), or `wide`
    const uint4* __restrict__ wleaf_ao;  // its leaves: first, count, reference leaf, triangle mask or 0
    const float4* __restrict__ wsub;     // group nodes inside whole leaves of `wide` (wleaf.w = node + 1)
    uint32_t n_nodes, n_cnodes;
};

__device__ inline V3 xyz(float4 v) { return mk(v.x, v.y, v.z); }

// RayAABB, Shaders/ao_step1.comp:43-52 (identical in every traversal shader).
__device__ inline float ray_aabb(V3 o, V3 d, float4 lo, float4 hi)
{
    float t1x = (lo.x - o.x) / d.x, t1y = (lo.y - o.y) / d.y, t1z = (lo.z - o.z) / d.z;
    float t2x = (hi.x - o.x) / d.x, t2y = (hi.y - o.y) / d.y, t2z = (hi.z - o.z) / d.z;
    float a = gmax(gmin(t1x, t2x), gmax(gmin(t1y, t2y), gmin(t1z, t2z)));
    float b = gmin(gmax(t1x, t2x), gmin(gmax(t1y, t2y), gmax(t1z, t2z)));
    return (b >= 0 && a <= b) ? a : FLT_MAX;
}


constexpr float kSlabRel = 1.0f / 2097152.0f;  // 2^-21 = 8 units in the last place
constexpr float kSlabAbs = 1e-37f;
constexpr float kSlowDir = 9.094947e-13f;  // 2^-40: smaller direction components use the exact test

struct Ray {
    V3 o, d, inv;
    V3 oi;  // RN(o * inv): the culling test's slab planes are fma(m, inv, -oi)
    bool slow;
};

// The reference's node test is RayAABB(o, d, box) < T with six IEEE
// divisions (Shaders/ao_step1.comp:43-52, :176).  Here the slab distances are
// first taken as (m - o) * RN(1/d), which lies within 3 units in the last
// place of RN((m - o)/d) (two roundings of a correctly rounded reciprocal and
// product against one of the quotient); max/min keep that relative bound.
// With an 8-ulp margin the outcome is decided whenever it cannot flip, and
// only the remaining boxes are retested with the reference's divisions, so
// the decision is always the reference's.
__device__ inline bool node_visit(const Ray& r, float4 lo, float4 hi, float T)
{
    if (!r.slow) {
        const float t1x = (lo.x - r.o.x) * r.inv.x, t2x = (hi.x - r.o.x) * r.inv.x;
        const float t1y = (lo.y - r.o.y) * r.inv.y, t2y = (hi.y - r.o.y) * r.inv.y;
        const float t1z = (lo.z - r.o.z) * r.inv.z, t2z = (hi.z - r.o.z) * r.inv.z;
        const float a = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
        const float b = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
        const float ea = __builtin_fmaf(fabsf(a), kSlabRel, kSlabAbs);
        const float eb = __builtin_fmaf(fabsf(b), kSlabRel, kSlabAbs);
        const float al = a - ea, ah = a + ea, bl = b - eb, bh = b + eb;
        if (bl >= 0.0f && ah <= bl && ah < T) return true;
        if (bh < 0.0f || al > bh || al >= T) return false;
    }
    return ray_aabb(r.o, r.d, lo, hi) < T;
}

// The node test of the watertight mapping: false only when the ray misses the
// box for certain, so that a ray through a vertex or edge reaches the leaves
// of every triangle sharing it.  The reference test rejects a box whose face
// lies in the plane a ray travels in: for d_i = 0 and o_i = lo_i or hi_i the
// slab distance is 0/0 = NaN and the comparisons fail.  Here such an axis is
// a containment test (lo_i <= o_i <= hi_i: the ray never leaves the plane),
// and the other slab distances carry the 8-ulp margin of node_visit (their
// error is at most 3 units in the last place), so entry <= exit, exit >= 0
// and entry < T hold whenever they hold for the exact distances.
__device__ inline bool node_visit_wt(const Ray& r, float4 lo, float4 hi, float T)
{
    float a, b;
    if (!r.slow) {
        const float t1x = (lo.x - r.o.x) * r.inv.x, t2x = (hi.x - r.o.x) * r.inv.x;
        const float t1y = (lo.y - r.o.y) * r.inv.y, t2y = (hi.y - r.o.y) * r.inv.y;
        const float t1z = (lo.z - r.o.z) * r.inv.z, t2z = (hi.z - r.o.z) * r.inv.z;
        a = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
        b = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
    } else {
        // components below 2^-40: the quotients themselves, zero components by containment
        a = -INFINITY;
        b = INFINITY;
        const float o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z};
        const float l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (d[i] == 0.0f) {
                if (o[i] < l[i] || o[i] > h[i]) return false;
            } else {
                const float t1 = (l[i] - o[i]) / d[i], t2 = (h[i] - o[i]) / d[i];
                a = fmaxf(a, fminf(t1, t2));
                b = fminf(b, fmaxf(t1, t2));
            }
        }
    }
    const float al = a - __builtin_fmaf(fabsf(a), kSlabRel, kSlabAbs);
    const float bh = b + __builtin_fmaf(fabsf(b), kSlabRel, kSlabAbs);
    return bh >= 0.0f && al <= bh && al < T;
}

// Conservative form of the same test for the culling tree: false only when
// the box is missed for certain.  The culling boxes contain the reference
// leaf boxes grown by pad = 2^-16 * max(1, max |coordinate|), which moves each
// slab plane of a reachable leaf by pad * |1/d_i| in t.  The slab distances
// here are fma(m, 1/d_i, -RN(o_i/d_i)) with |m|, |o_i| <= max |coordinate|
// (+pad): their error against (m - o_i)/d_i is a few units in the last place
// of max(|m|, |o_i|) * |1/d_i|, i.e. below 2^-21 * max(1, max |coordinate|) *
// |1/d_i|, far inside the padding; the reference's own rounding of its
// quotients is of the same order.  So a culling box is never rejected when a
// reference leaf inside it passes the reference test, and no margin is
// needed.  `al` (entry) orders children and prunes thickness entries.
__device__ inline bool box_maybe(const Ray& r, float4 lo, float4 hi, float T, float& al)
{
    const float t1x = __builtin_fmaf(lo.x, r.inv.x, -r.oi.x), t2x = __builtin_fmaf(hi.x, r.inv.x, -r.oi.x);
    const float t1y = __builtin_fmaf(lo.y, r.inv.y, -r.oi.y), t2y = __builtin_fmaf(hi.y, r.inv.y, -r.oi.y);
    const float t1z = __builtin_fmaf(lo.z, r.inv.z, -r.oi.z), t2z = __builtin_fmaf(hi.z, r.inv.z, -r.oi.z);
    const float a = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
    const float b = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
    al = a;
    return a <= b && b >= 0.0f && a < T;
}

__device__ inline bool node_maybe(const Ray& r, float4 lo, float4 hi, float T)
{
    float al;
    return box_maybe(r, lo, hi, T, al);
}

// AO and BN decide a ray on its first qualifying hit (ao_step1.comp:158,
// bentnormals_step1.comp:170), and the reference finds that hit iff the
// triangle's leaf is reachable: every node from the root down to it has slab
// entry < maxD (before the first hit the running minimum is FLT_MAX, and once
// a hit below maxD exists the ray is decided).  So a hit found through the
// culling tree counts exactly when its reference leaf's chain passes the
// reference test.
//
// Only the leaf and the ancestors its reachability links reach are tested:
// when a parent's box P contains its child's box C, every rounded step of
// the test is monotone -- RN(RN(m - o) / d) in the plane m for a fixed d != 0,
// then min, max -- so P's slab interval contains C's and P passes whenever C
// does (entry a_P <= a_C < T, exit b_P >= b_C >= max(0, a_C)).  The rays that
// come here have every |d_i| >= 2^-40, so no quotient is NaN.
__device__ inline bool chain_ok(const Ray& r, const DevScene& sc, uint32_t leaf, float T)
{
    for (uint32_t x = leaf; x != kMiss; x = sc.chain[x])
        if (!node_visit(r, sc.nodes[2 * x], sc.nodes[2 * x + 1], T)) return false;
    return true;
}

// Shaders/ao_step1.comp:109-147 with the per-triangle terms read from `tri`.
// Returns t, or FLT_MAX.  Triangles whose t falls outside [tmin, tmax) cannot
// change the caller's result and are rejected before the barycentrics.
// Second stage of the triangle test: t = num / nd (num = a.n - o.n, nd =
// d.n, both as the shader computes them), its range, then the barycentrics.
__device__ inline float hemi_tri_t(V3 o, V3 d, float num, float nd, uint32_t k, const DevScene& sc, float tmin,
                                   float tmax)
{
    const float t = num / nd;
    if (!(t >= 0 && t >= tmin && t < tmax)) return FLT_MAX;
    const float4 q1 = sc.trest[3 * k], q2 = sc.trest[3 * k + 1], q3 = sc.trest[3 * k + 2];
    const V3 p = o + d * t;
    const V3 v2 = p - xyz(q1), v0 = xyz(q2), v1 = xyz(q3);
    const float d00 = q1.w, d01 = q2.w, d11 = q3.w;
    const float d20 = dot(v2, v0), d21 = dot(v2, v1);
    const float denom = d00 * d11 - d01 * d01;
    const float y = (d11 * d20 - d01 * d21) / denom;
    const float z = (d00 * d21 - d01 * d20) / denom;
    const float x = 1.0f - y - z;
    return (x >= 0 && y >= 0 && y <= 1 && z >= 0 && z <= 1) ? t : FLT_MAX;
}

__device__ inline float hemi_tri_q(V3 o, V3 d, float4 q0, uint32_t k, const DevScene& sc, float tmin,
                                   float tmax)
{
    const V3 n = xyz(q0);
    const float nd = dot(d, n);
    if (!(fabsf(nd) > 0)) return FLT_MAX;
    const float pn = dot(o, n);
    return hemi_tri_t(o, d, q0.w - pn, nd, k, sc, tmin, tmax);
}

__device__ inline float hemi_tri(V3 o, V3 d, uint32_t k, const DevScene& sc, float tmin, float tmax)
{
    return hemi_tri_q(o, d, sc.tri[k], k, sc, tmin, tmax);
}

struct Basis {
    V3 o, d, tx, ty;
};

// Shaders/ao_step0.comp:39-56 for a texel whose mapping hit soup vertex t.
__device__ inline Basis ray_basis(const DevScene& sc, float4 c, uint32_t t)
{
    Basis B;
    const V3 p0 = xyz(sc.spos[t]), p1 = xyz(sc.spos[t + 1]), p2 = xyz(sc.spos[t + 2]);
    B.o = p0 * c.y + p1 * c.z + p2 * c.w;
    const V3 n0 = xyz(sc.snrm[t]), n1 = xyz(sc.snrm[t + 1]), n2 = xyz(sc.snrm[t + 2]);
    B.d = normalize(n0 * c.y + n1 * c.z + n2 * c.w);
    const V3 d = B.d;
    B.ty = normalize(fabsf(d.x) > fabsf(d.y) ? mk(d.z, 0, -d.x) : mk(0, d.z, -d.y));
    B.tx = cross(d, B.ty);
    return B;
}

// Sample direction of ao_step1.comp:153-155 (thick_step1.comp:144-150 flips d).
__device__ inline V3 sample_dir(const Basis& B, V3 d, float4 rs)
{
    return normalize(B.tx * rs.x + B.ty * rs.y + d * rs.z);
}

}  // namespace
}  // namespace fornos
