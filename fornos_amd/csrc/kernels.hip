// kernels.hip -- gfx950 kernels of the baking hot path and the device side of
// the C ABI (include/fornos_hip.h).
//
// HBM layout of a scene (all soup-ordered, i.e. in the DFS leaf order of the
// reference flatten, Src/meshmapping.cpp:99-116):
//   nodes  2 x float4 per node: (min.xyz, first soup triangle),
//                               (max.xyz, jump | kLeafFlag+count)
//   tri    4 x float4 per triangle: the per-triangle terms of the hemisphere
//          and mapping predicates, computed once on the host with the
//          shader's own operation order:
//          (n.xyz, dot(a,n)), (a.xyz, d00), (v0.xyz, d01), (v1.xyz, d11)
//          with n = normalize(cross(b-a, c-a)), v0 = b-a, v1 = c-a
//   spos   3 x float4 per triangle (reference Vector4 soup positions)
//   snrm   3 x float4 per triangle (reference Vector4 soup normals)
// Texel data: pix 2 x float4 per texel (p, d), pixt 3 x float4 (n, t, b),
// coords float4 (t, b0, b1, b2), tidx uint (soup vertex index).
//
// Everything is compiled with -ffp-contract=off and IEEE-correct fp32
// division/sqrt so that results are bit-identical to the CPU restatement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "host.h"

#pragma clang fp contract(off)

using namespace fornos;

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
    const uint32_t* __restrict__ parent;  // reference tree parent links
    const float4* __restrict__ cnodes;   // culling tree over the reference leaves
    const uint32_t* __restrict__ cref;   // reference leaf of each culling leaf
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
// slab entry is < the running best, else jump past its subtree.
template <int MODE>
__device__ inline void map_pass(V3 o, V3 d, const DevScene& sc, float& cur, uint32_t& idx, V3& bc)
{
    uint32_t i = 0;
    while (i < sc.n_nodes) {
        const float4 lo = sc.nodes[2 * i], hi = sc.nodes[2 * i + 1];
        const uint32_t w = __float_as_uint(hi.w);
        if (ray_aabb(o, d, lo, hi) < cur) {
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

template <bool CULL>
__global__ __launch_bounds__(kBlock) void k_meshmap(const float4* __restrict__ pix, uint32_t begin,
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
    if (CULL) {
        // meshmapping_nobackfaces.comp:182-199: one running t shared by both passes
        map_pass<1>(p, d, sc, t, idx, bc);
        map_pass<2>(p, -d, sc, t, idx, bc);
    } else {
        // meshmapping.comp:139-162: the backward pass starts from the forward t
        map_pass<0>(p, d, sc, t, idx, bc);
        map_pass<0>(p, -d, sc, t, idx, bc);
    }
    coords[g] = make_float4(t, bc.x, bc.y, bc.z);
    tidx[g] = idx;
}

// ---------------------------------------------------------------------------
// Hemisphere passes (AO, bent normals, thickness).
//
// k_rays: persistent waves.  Each lane owns one ray (ao_step0 ray basis +
// ao/bentnormals/thick step1 traversal); a lane whose ray has finished takes
// a new one from a global counter, one atomic per wave for all its idle
// lanes, so a wave never idles behind its slowest ray.  Every ray writes one
// result (a hit byte, or t for thickness) to a per-launch buffer; k_reduce
// then folds each texel's S results in sample order (step2 + tangentspace),
// the same split as the reference's step1/step2 dispatches.
// ---------------------------------------------------------------------------
enum { kAO = 0, kBN = 1, kThick = 2 };

constexpr uint32_t kRefill = 16;           // refill a wave once this many lanes are idle
constexpr float kSlabRel = 1.0f / 2097152.0f;  // 2^-21 = 8 units in the last place
constexpr float kSlabAbs = 1e-37f;
constexpr float kSlowDir = 9.094947e-13f;  // 2^-40: smaller direction components use the exact test

struct Ray {
    V3 o, d, inv;
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

// Conservative form of the same test for the culling tree: false only when
// the box is missed for certain (its boxes are padded, so certainty suffices).
__device__ inline bool node_maybe(const Ray& r, float4 lo, float4 hi, float T)
{
    const float t1x = (lo.x - r.o.x) * r.inv.x, t2x = (hi.x - r.o.x) * r.inv.x;
    const float t1y = (lo.y - r.o.y) * r.inv.y, t2y = (hi.y - r.o.y) * r.inv.y;
    const float t1z = (lo.z - r.o.z) * r.inv.z, t2z = (hi.z - r.o.z) * r.inv.z;
    const float a = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
    const float b = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
    const float al = a - __builtin_fmaf(fabsf(a), kSlabRel, kSlabAbs);
    const float bh = b + __builtin_fmaf(fabsf(b), kSlabRel, kSlabAbs);
    return !(bh < 0.0f || al > bh || al >= T);
}

// AO and BN decide a ray on its first qualifying hit (ao_step1.comp:158,
// bentnormals_step1.comp:170), and the reference finds that hit iff the
// triangle's leaf is reachable: every node from the root down to it has slab
// entry < maxD (before the first hit the running minimum is FLT_MAX, and once
// a hit below maxD exists the ray is decided).  So a hit found through the
// culling tree counts exactly when its reference leaf's chain passes the
// reference test.
__device__ inline bool chain_ok(const Ray& r, const DevScene& sc, uint32_t leaf, float T)
{
    for (uint32_t x = leaf; x != kMiss; x = sc.parent[x])
        if (!node_visit(r, sc.nodes[2 * x], sc.nodes[2 * x + 1], T)) return false;
    return true;
}

// Shaders/ao_step1.comp:109-147 with the per-triangle terms read from `tri`.
// Returns t, or FLT_MAX.  Triangles whose t falls outside [tmin, tmax) cannot
// change the caller's result and are rejected before the barycentrics.
__device__ inline float hemi_tri_q(V3 o, V3 d, float4 q0, uint32_t k, const DevScene& sc, float tmin,
                                   float tmax)
{
    const V3 n = xyz(q0);
    const float nd = dot(d, n);
    if (!(fabsf(nd) > 0)) return FLT_MAX;
    const float pn = dot(o, n);
    const float t = (q0.w - pn) / nd;
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

struct RayArgs {
    DevScene sc;
    const float4* __restrict__ coords;
    const uint32_t* __restrict__ tidx;
    const float4* __restrict__ samples;
    uint8_t* __restrict__ hit;          // AO / BN: 1 when the ray hit (per ray of the launch)
    float* __restrict__ tval;           // thickness: closest t, FLT_MAX when nothing was hit
    unsigned int* __restrict__ counter;  // next ray of the launch to hand out
    unsigned long long* __restrict__ stats;  // node tests, triangle tests, rays (COUNT only)
    uint32_t S, P, t0, ntex;            // texels [t0, t0 + ntex), S rays each
    float mind, maxd;
};

// One step of raycastBVH (Shaders/ao_step1.comp:166-193) in the reference's
// stackless order.  AO: any hit with t in [minD, maxD) decides the ray
// (ao_step1.comp:158); nodes are visited iff entry < maxD (the running minimum
// only prunes nodes the maxD test already prunes).  BN: any hit with
// t >= minD decides it (bentnormals_step1.comp:170).  Thickness: closest hit
// with t >= minD over nodes with entry < min(best, maxD).
template <int KIND, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_rays(RayArgs A)
{
    const uint32_t total = A.ntex * A.S;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t n_nodes = A.sc.n_nodes;
    Ray ray;
    ray.o = ray.d = ray.inv = mk(0, 0, 0);
    ray.slow = false;
    uint32_t rid = 0, node = 0;
    float best = FLT_MAX;
    bool active = false, exhausted = false, ref_path = false;
    uint32_t cn = 0, ct = 0, nr = 0;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        const uint32_t nidle = __popcll(idle);
        if (!exhausted && (nidle >= kRefill || nidle == 64)) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(A.counter, nidle);
            base = __shfl(base, 0);
            if (base + nidle >= total) exhausted = true;
            if (!active) {
                const uint32_t r = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1));
                if (r < total) {
                    const uint32_t q = r / A.S, s = r - q * A.S, g = A.t0 + q;
                    const uint32_t t = A.tidx[g];
                    if (t == kMiss) {
                        if (KIND == kThick) A.tval[r] = FLT_MAX;
                        else A.hit[r] = 0;
                    } else {
                        const Basis B = ray_basis(A.sc, A.coords[g], t);
                        const V3 d = KIND == kThick ? -B.d : B.d;
                        ray.o = B.o;
                        ray.d = sample_dir(B, d, A.samples[(size_t)(g % A.P) * A.S + s]);
                        ray.inv = mk(1.0f / ray.d.x, 1.0f / ray.d.y, 1.0f / ray.d.z);
                        ray.slow = !(fabsf(ray.d.x) >= kSlowDir && fabsf(ray.d.y) >= kSlowDir &&
                                     fabsf(ray.d.z) >= kSlowDir);
                        ref_path = ray.slow;  // tiny direction components: reference tree, exact test
                        rid = r;
                        node = 0;
                        best = FLT_MAX;
                        active = true;
                        if (COUNT) ++nr;
                    }
                }
            }
        }
        if (exhausted && __ballot(active) == 0) break;
        if (active && !ref_path) {
            // The culling tree.  AO / BN: a qualifying hit decides the ray iff
            // its reference leaf is reachable (chain_ok with maxD).
            // Thickness: a hit at t < best is taken when its leaf's chain
            // passes with min(maxD, t): then the reference, whose running
            // minimum never drops below the final one, visits that leaf too,
            // and the minimum over reachable leaves is the reference's value.
            // A leaf reachable with maxD but not with min(maxD, t) (a box
            // entered no earlier than the hit) sends the ray to the reference
            // tree instead.
            const float4 lo = A.sc.cnodes[2 * node], hi = A.sc.cnodes[2 * node + 1];
            const uint32_t w = __float_as_uint(hi.w);
            const float T = KIND == kThick ? fminf(best, A.maxd) : A.maxd;
            bool found = false, fallback = false;
            if (COUNT) ++cn;
            if (node_maybe(ray, lo, hi, T)) {
                if (w & kLeafFlag) {
                    const uint32_t first = __float_as_uint(lo.w), end = first + (w & ~kLeafFlag);
                    const float tri_hi = KIND == kAO ? A.maxd : INFINITY;
                    bool stop = false;
                    for (uint32_t k0 = first; k0 < end && !stop; k0 += 4) {
                        // the leaf's (n, a.n) records are contiguous: fetch four at once
                        float4 q[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) q[j] = k0 + j < end ? A.sc.tri[k0 + j] : make_float4(0, 0, 0, 0);
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const uint32_t k = k0 + j;
                            if (stop || k >= end) continue;
                            if (COUNT) ++ct;
                            const float t = hemi_tri_q(ray.o, ray.d, q[j], k, A.sc, A.mind,
                                                       KIND == kThick ? best : tri_hi);
                            if (t == FLT_MAX) continue;
                            const uint32_t leaf = A.sc.cref[node];
                            if (KIND != kThick) {
                                found = chain_ok(ray, A.sc, leaf, A.maxd);
                                stop = true;  // the other triangles of the leaf share its chain
                            } else if (chain_ok(ray, A.sc, leaf, fminf(A.maxd, t))) {
                                best = t;
                            } else {
                                fallback = chain_ok(ray, A.sc, leaf, A.maxd);
                                stop = true;  // unreachable leaf: none of its triangles count
                            }
                        }
                    }
                }
                ++node;
            } else {
                node = (w & kLeafFlag) ? node + 1 : w;
            }
            if (fallback) {
                ref_path = true;  // redo this ray on the reference tree
                node = 0;
                best = FLT_MAX;
            } else if (found || node >= A.sc.n_cnodes) {
                if (KIND == kThick) A.tval[rid] = best;
                else A.hit[rid] = found ? 1 : 0;
                active = false;
            }
        } else if (active) {
            const float4 lo = A.sc.nodes[2 * node], hi = A.sc.nodes[2 * node + 1];
            const uint32_t w = __float_as_uint(hi.w);
            const float T = KIND == kThick ? fminf(best, A.maxd) : A.maxd;
            bool done = false, found = false;
            if (COUNT) ++cn;
            if (node_visit(ray, lo, hi, T)) {
                if (w & kLeafFlag) {
                    const uint32_t first = __float_as_uint(lo.w), end = first + (w & ~kLeafFlag);
                    const float tri_hi = KIND == kAO ? A.maxd : (KIND == kBN ? INFINITY : best);
                    for (uint32_t k = first; k < end; ++k) {
                        if (COUNT) ++ct;
                        const float t = hemi_tri(ray.o, ray.d, k, A.sc, A.mind, KIND == kThick ? best : tri_hi);
                        if (t != FLT_MAX) {
                            if (KIND != kThick) { found = true; break; }
                            best = t;
                        }
                    }
                }
                ++node;
            } else {
                node = (w & kLeafFlag) ? node + 1 : w;
            }
            if (found || node >= n_nodes) done = true;
            if (done) {
                if (KIND == kThick) A.tval[rid] = best;
                else A.hit[rid] = found ? 1 : 0;
                active = false;
            }
        }
    }
    if (COUNT) {
        unsigned long long a = cn, b = ct, c = nr;
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_down(a, off);
            b += __shfl_down(b, off);
            c += __shfl_down(c, off);
        }
        if (lane == 0) {
            atomicAdd(&A.stats[0], a);
            atomicAdd(&A.stats[1], b);
            atomicAdd(&A.stats[2], c);
        }
    }
}

struct ReduceArgs {
    DevScene sc;
    const float4* __restrict__ coords;
    const uint32_t* __restrict__ tidx;
    const float4* __restrict__ samples;
    const float4* __restrict__ pixt;    // n, t, b per texel (BN tangent space) or null
    const uint8_t* __restrict__ hit;
    const float* __restrict__ tval;
    float* __restrict__ out;            // 1 or 3 floats per texel (global texel index)
    uint32_t* __restrict__ counts;      // per texel
    uint32_t S, P, t0, ntex;
    float maxd;
    int tangent;
};

// step2 reductions in sample order (ao_step2.comp:24-29, thick_step2.comp:24-29,
// bentnormals_step2.comp:25-34) and tangentspace.comp:22-41.  Texels whose
// mapping missed get AO 1, thickness maxD and bent normal (0,0,0).
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_reduce(ReduceArgs A)
{
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= A.ntex) return;
    const uint32_t g = A.t0 + q, S = A.S;
    const size_t base = (size_t)q * S;
    const uint32_t t = A.tidx[g];
    uint32_t cnt = 0;
    if (KIND == kAO) {
        if (t != kMiss)
            for (uint32_t s = 0; s < S; ++s) cnt += A.hit[base + s];
        A.out[g] = t == kMiss ? 1.0f : 1.0f - (float)cnt / (float)S;
    } else if (KIND == kThick) {
        float acc = 0;
        if (t != kMiss)
            for (uint32_t s = 0; s < S; ++s) {
                const float v = A.tval[base + s];
                if (v != FLT_MAX) ++cnt;
                acc += v != FLT_MAX ? v : A.maxd;
            }
        A.out[g] = t == kMiss ? A.maxd : acc / (float)S;
    } else {
        V3 r = mk(0, 0, 0);
        if (t != kMiss) {
            const Basis B = ray_basis(A.sc, A.coords[g], t);
            const float4* smp = A.samples + (size_t)(g % A.P) * S;
            V3 acc = mk(0, 0, 0);
            for (uint32_t s = 0; s < S; ++s) {
                if (A.hit[base + s]) { ++cnt; continue; }   // adds vec3(0): acc never holds -0
                acc = acc + sample_dir(B, B.d, smp[s]);
            }
            r = normalize(acc);
            if (A.tangent) {
                const V3 n = xyz(A.pixt[3 * g]), tt = xyz(A.pixt[3 * g + 1]), b = xyz(A.pixt[3 * g + 2]);
                const V3 d0 = mk(n.z * b.y - n.y * b.z, n.x * b.z - n.z * b.x, n.y * b.x - n.x * b.y);
                const V3 d1 = mk(tt.z * n.y - tt.y * n.z, tt.x * n.z - n.x * tt.z, n.x * tt.y - tt.x * n.y);
                const V3 d2 = mk(tt.y * b.z - tt.z * b.y, tt.z * b.x - tt.x * b.z, tt.x * b.y - tt.y * b.x);
                r = normalize(mk(dot(r, d0), dot(r, d1), dot(r, d2)));
            }
        }
        A.out[3 * g] = r.x;
        A.out[3 * g + 1] = r.y;
        A.out[3 * g + 2] = r.z;
    }
    A.counts[g] = cnt;
}

// Clears n 32-bit words (a kernel rather than hipMemsetAsync, so the whole
// pass can be captured into a HIP graph).
__global__ void k_zero(uint32_t* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
// ---------------------------------------------------------------------------
// Texel passes: heights.comp, positions.comp, normals.comp (+ tangentspace).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_height(const float4* __restrict__ coords, uint32_t begin,
                                                    uint32_t end, float* __restrict__ out)
{
    const uint32_t g = begin + blockIdx.x * kBlock + threadIdx.x;
    if (g >= end) return;
    const float h = coords[g].x;
    out[g] = h != FLT_MAX ? h : 0.0f;
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
    const uint32_t t = tidx[g];
    V3 r = mk(0, 0, 0);
    if (t != kMiss) {
        const float4 c = coords[g];
        r = xyz(soup[t]) * c.y + xyz(soup[t + 1]) * c.z + xyz(soup[t + 2]) * c.w;
        if (NORMALS) {
            r = normalize(r);
            if (tangent) {
                const V3 n = xyz(pixt[3 * g]), tt = xyz(pixt[3 * g + 1]), b = xyz(pixt[3 * g + 2]);
                const V3 d0 = mk(n.z * b.y - n.y * b.z, n.x * b.z - n.z * b.x, n.y * b.x - n.x * b.y);
                const V3 d1 = mk(tt.z * n.y - tt.y * n.z, tt.x * n.z - n.x * tt.z, n.x * tt.y - tt.x * n.y);
                const V3 d2 = mk(tt.y * b.z - tt.z * b.y, tt.z * b.x - tt.x * b.z, tt.x * b.y - tt.y * b.x);
                r = normalize(mk(dot(r, d0), dot(r, d1), dot(r, d2)));
            }
        }
    }
    out[3 * g] = r.x;
    out[3 * g + 1] = r.y;
    out[3 * g + 2] = r.z;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            return set_error(e_ == hipErrorOutOfMemory ? FORNOS_OOM : FORNOS_HIP_ERROR,   \
                             std::string(#expr) + ": " + hipGetErrorString(e_));          \
        }                                                                                 \
    } while (0)

template <typename T>
fornos_status dev_alloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return set_error(e == hipErrorOutOfMemory ? FORNOS_OOM : FORNOS_HIP_ERROR,
                         std::string("hipMalloc(") + std::to_string(sizeof(T) * count) +
                             " bytes): " + hipGetErrorString(e));
    }
    return FORNOS_OK;
}

void dev_free(void* p)
{
    if (p) (void)hipFree(p);
}

fornos_status require_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return set_error(FORNOS_UNSUPPORTED, "no HIP device is visible");
    }
    return FORNOS_OK;
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit Timer(hipStream_t st) : s(st) {}
    fornos_status start()
    {
        HIP_TRY(hipEventCreate(&a));
        HIP_TRY(hipEventCreate(&b));
        HIP_TRY(hipEventRecord(a, s));
        return FORNOS_OK;
    }
    fornos_status stop(float* ms)
    {
        HIP_TRY(hipEventRecord(b, s));
        HIP_TRY(hipEventSynchronize(b));
        float v = 0;
        HIP_TRY(hipEventElapsedTime(&v, a, b));
        if (ms) *ms = v;
        return FORNOS_OK;
    }
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

struct fornos_scene {
    int device = 0;
    uint32_t persistent_blocks = 2048;  // k_rays grid: 8 blocks of 256 lanes per CU
    hipStream_t stream = nullptr;
    DevScene d{};
    float4 *nodes = nullptr, *tri = nullptr, *spos = nullptr, *snrm = nullptr, *cnodes = nullptr;
    uint32_t *parent = nullptr, *cref = nullptr;
    FlatBVH bvh;
    fornos_scene_info info{};
    ~fornos_scene()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        dev_free(nodes); dev_free(tri); dev_free(spos); dev_free(snrm);
        dev_free(cnodes); dev_free(parent); dev_free(cref);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct fornos_meshmap {
    fornos_scene* sc = nullptr;
    uint32_t n = 0, width = 0, height = 0;
    uint32_t work_count = 0;  // n rounded up to 64, Src/meshmapping.cpp:165
    uint32_t offset = 0;      // texels mapped so far
    int cull = 1;
    float4 *pix = nullptr, *pixt = nullptr, *coords = nullptr;
    uint32_t* tidx = nullptr;
    ~fornos_meshmap()
    {
        if (sc && sc->stream) (void)hipStreamSynchronize(sc->stream);
        dev_free(pix); dev_free(pixt); dev_free(coords); dev_free(tidx);
    }
};

struct fornos_solver {
    int kind = 0;
    fornos_solver_params p{};
    fornos_meshmap* mm = nullptr;
    uint32_t P = 0;
    uint64_t work_total = 0, offset = 0;  // work items (rays or texels)
    float4* samples = nullptr;
    float* out = nullptr;
    uint32_t* counts = nullptr;
    unsigned long long* stats = nullptr;
    void* raybuf = nullptr;            // per-ray results of one launch (hit bytes or t)
    uint32_t chunk_texels = 0;         // texels per k_rays launch
    uint32_t* counter = nullptr;       // k_rays work counter
    // per-launch events of the last timed run: (before k_rays, after k_rays, after k_reduce)
    std::vector<hipEvent_t> events;
    bool timed = false;
    float rays_ms = 0, reduce_ms = 0;
    uint32_t launches = 0;
    ~fornos_solver()
    {
        if (mm && mm->sc && mm->sc->stream) (void)hipStreamSynchronize(mm->sc->stream);
        dev_free(samples); dev_free(out); dev_free(counts); dev_free(stats);
        dev_free(raybuf); dev_free(counter);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
    bool hemi() const { return kind >= FORNOS_SOLVER_AO; }
    uint32_t comps() const
    {
        return (kind == FORNOS_SOLVER_POSITION || kind == FORNOS_SOLVER_NORMALS || kind == FORNOS_SOLVER_BENT_NORMALS) ? 3 : 1;
    }
};

namespace {

// Launch a hemisphere kernel over texels [t0, t1).
template <int KIND>
fornos_status launch_hemi(fornos_solver* s, uint32_t t0, uint32_t t1, bool count)
{
    fornos_meshmap* mm = s->mm;
    fornos_scene* sc = mm->sc;
    const uint32_t S = s->p.sample_count;
    RayArgs A;
    A.sc = sc->d;
    A.coords = mm->coords;
    A.tidx = mm->tidx;
    A.samples = s->samples;
    A.hit = KIND == kThick ? nullptr : (uint8_t*)s->raybuf;
    A.tval = KIND == kThick ? (float*)s->raybuf : nullptr;
    A.counter = s->counter;
    A.stats = s->stats;
    A.S = S;
    A.P = s->P;
    A.mind = s->p.min_distance;
    A.maxd = s->p.max_distance;
    ReduceArgs R;
    R.sc = sc->d;
    R.coords = mm->coords;
    R.tidx = mm->tidx;
    R.samples = s->samples;
    R.pixt = mm->pixt;
    R.hit = A.hit;
    R.tval = A.tval;
    R.out = s->out;
    R.counts = s->counts;
    R.S = S;
    R.P = s->P;
    R.maxd = s->p.max_distance;
    R.tangent = s->p.tangent_space;
    if (count) hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, sc->stream, (uint32_t*)s->stats, 6u);
    for (uint32_t c0 = t0; c0 < t1; c0 += s->chunk_texels) {
        const uint32_t nt = std::min(s->chunk_texels, t1 - c0);
        A.t0 = R.t0 = c0;
        A.ntex = R.ntex = nt;
        const uint32_t blocks = std::min<uint32_t>(sc->persistent_blocks, div_up((uint64_t)nt * S, kBlock));
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, sc->stream, s->counter, 1u);
        hipEvent_t* ev = nullptr;
        if (s->timed) {
            s->events.resize(s->events.size() + 3, nullptr);
            ev = &s->events[s->events.size() - 3];
            for (int i = 0; i < 3; ++i) HIP_TRY(hipEventCreate(&ev[i]));
            HIP_TRY(hipEventRecord(ev[0], sc->stream));
        }
        if (count)
            hipLaunchKernelGGL((k_rays<KIND, true>), dim3(blocks), dim3(kBlock), 0, sc->stream, A);
        else
            hipLaunchKernelGGL((k_rays<KIND, false>), dim3(blocks), dim3(kBlock), 0, sc->stream, A);
        if (ev) HIP_TRY(hipEventRecord(ev[1], sc->stream));
        hipLaunchKernelGGL(k_reduce<KIND>, dim3(div_up(nt, kBlock)), dim3(kBlock), 0, sc->stream, R);
        if (ev) HIP_TRY(hipEventRecord(ev[2], sc->stream));
        HIP_TRY(hipGetLastError());
    }
    return FORNOS_OK;
}

fornos_status solver_launch(fornos_solver* s, uint32_t t0, uint32_t t1, bool count)
{
    fornos_meshmap* mm = s->mm;
    fornos_scene* sc = mm->sc;
    if (t1 <= t0) return FORNOS_OK;
    const uint32_t blocks = div_up(t1 - t0, kBlock);
    switch (s->kind) {
    case FORNOS_SOLVER_HEIGHT:
        hipLaunchKernelGGL(k_height, dim3(blocks), dim3(kBlock), 0, sc->stream, mm->coords, t0, t1, s->out);
        break;
    case FORNOS_SOLVER_POSITION:
        hipLaunchKernelGGL(k_attr<false>, dim3(blocks), dim3(kBlock), 0, sc->stream, mm->coords, mm->tidx,
                           sc->spos, mm->pixt, t0, t1, 0, s->out);
        break;
    case FORNOS_SOLVER_NORMALS:
        hipLaunchKernelGGL(k_attr<true>, dim3(blocks), dim3(kBlock), 0, sc->stream, mm->coords, mm->tidx,
                           sc->snrm, mm->pixt, t0, t1, s->p.tangent_space, s->out);
        break;
    case FORNOS_SOLVER_AO: return launch_hemi<kAO>(s, t0, t1, count);
    case FORNOS_SOLVER_BENT_NORMALS: return launch_hemi<kBN>(s, t0, t1, count);
    case FORNOS_SOLVER_THICKNESS: return launch_hemi<kThick>(s, t0, t1, count);
    }
    HIP_TRY(hipGetLastError());
    return FORNOS_OK;
}

}  // namespace

extern "C" {

const char* fornos_version(void) { return "fornos-mi355x 0.1 (gfx950)"; }

int fornos_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

fornos_status fornos_scene_create(const fornos_mesh_view* hi, const fornos_bvh_params* params,
                                  fornos_scene** out)
{
    if (!hi || !out) return set_error(FORNOS_INVALID_ARG, "fornos_scene_create: null argument");
    *out = nullptr;
    fornos_bvh_params bp = params ? *params : fornos_bvh_params{FORNOS_BVH_REFERENCE, 8, 8192};
    if (bp.layout != FORNOS_BVH_REFERENCE)
        return set_error(FORNOS_UNSUPPORTED, "fornos_scene_create: unknown BVH layout");
    if (bp.max_tris_per_leaf == 0) return set_error(FORNOS_INVALID_ARG, "fornos_scene_create: max_tris_per_leaf must be >= 1");
    if (hi->n_triangles == 0) return set_error(FORNOS_INVALID_ARG, "fornos_scene_create: the hi-poly mesh has no triangles");
    if (!hi->normals) return set_error(FORNOS_INVALID_ARG, "fornos_scene_create: the hi-poly mesh has no normals");
    if ((uint64_t)hi->n_triangles * 3 >= kMiss) return set_error(FORNOS_UNSUPPORTED, "fornos_scene_create: too many triangles");
    if (fornos_status st = require_device()) return st;

    fornos_mesh* m = nullptr;
    if (fornos_status st = fornos_mesh_from_view(hi, &m)) return st;
    std::unique_ptr<fornos_mesh> mesh(m);
    for (uint32_t v = 0; v < mesh->nv(); ++v)
        if (mesh->vertices[3 * v + 2] == UINT32_MAX)
            return set_error(FORNOS_INVALID_ARG, "fornos_scene_create: a hi-poly vertex has no normal");

    std::unique_ptr<fornos_scene> sc(new fornos_scene());
    HIP_TRY(hipGetDevice(&sc->device));
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, sc->device));
    sc->persistent_blocks = (uint32_t)std::max(1, cus) * 8;
    const auto t_begin = std::chrono::steady_clock::now();
    build_reference_bvh(*mesh, bp.max_tris_per_leaf, bp.max_depth, sc->bvh);
    const FlatBVH& B = sc->bvh;
    const uint32_t nt = (uint32_t)B.soup_tri.size();
    std::vector<float4> tri((size_t)4 * nt), spos((size_t)3 * nt), snrm((size_t)3 * nt);
    for (uint32_t k = 0; k < nt; ++k) {
        const uint32_t t = B.soup_tri[k];
        V3 p[3], n[3];
        for (int c = 0; c < 3; ++c) {
            const uint32_t v = mesh->triangles[3 * t + c];
            p[c] = mesh->positions[mesh->vertices[3 * v]];
            n[c] = mesh->normals[mesh->vertices[3 * v + 2]];
            spos[3 * k + c] = make_float4(p[c].x, p[c].y, p[c].z, 0.0f);
            snrm[3 * k + c] = make_float4(n[c].x, n[c].y, n[c].z, 0.0f);
        }
        const V3 v0 = p[1] - p[0], v1 = p[2] - p[0];
        const V3 nn = normalize(cross(v0, v1));
        tri[k] = make_float4(nn.x, nn.y, nn.z, dot(p[0], nn));
        tri[nt + 3 * (size_t)k + 0] = make_float4(p[0].x, p[0].y, p[0].z, dot(v0, v0));
        tri[nt + 3 * (size_t)k + 1] = make_float4(v0.x, v0.y, v0.z, dot(v0, v1));
        tri[nt + 3 * (size_t)k + 2] = make_float4(v1.x, v1.y, v1.z, dot(v1, v1));
    }
    // culling tree over the reference leaves, boxes padded by 2^-16 of the
    // largest coordinate magnitude (far above the predicate's rounding)
    float amax = 1.0f;
    for (const float4& q : spos) amax = std::max(amax, std::max(fabsf(q.x), std::max(fabsf(q.y), fabsf(q.z))));
    CullBVH cull;
    {
        std::vector<float> sp4((size_t)4 * spos.size());
        std::memcpy(sp4.data(), spos.data(), sizeof(float) * sp4.size());
        build_cull_bvh(B, sp4, amax * (1.0f / 65536.0f), cull);
    }
    const std::vector<uint32_t> parent = reference_parents(B);
    const double build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count();

    HIP_TRY(hipStreamCreateWithFlags(&sc->stream, hipStreamNonBlocking));
    if (fornos_status st = dev_alloc(&sc->cnodes, (size_t)2 * cull.n_nodes)) return st;
    if (fornos_status st = dev_alloc(&sc->cref, cull.ref_leaf.size())) return st;
    if (fornos_status st = dev_alloc(&sc->parent, parent.size())) return st;
    HIP_TRY(hipMemcpy(sc->cnodes, cull.nodes.data(), sizeof(uint32_t) * 8 * (size_t)cull.n_nodes, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(sc->cref, cull.ref_leaf.data(), sizeof(uint32_t) * cull.ref_leaf.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(sc->parent, parent.data(), sizeof(uint32_t) * parent.size(), hipMemcpyHostToDevice));
    sc->d.cnodes = sc->cnodes;
    sc->d.cref = sc->cref;
    sc->d.parent = sc->parent;
    sc->d.n_cnodes = cull.n_nodes;
    if (fornos_status st = dev_alloc(&sc->nodes, (size_t)2 * B.n_nodes)) return st;
    if (fornos_status st = dev_alloc(&sc->tri, tri.size())) return st;
    if (fornos_status st = dev_alloc(&sc->spos, spos.size())) return st;
    if (fornos_status st = dev_alloc(&sc->snrm, snrm.size())) return st;
    HIP_TRY(hipMemcpy(sc->nodes, B.nodes.data(), sizeof(uint32_t) * 8 * (size_t)B.n_nodes, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(sc->tri, tri.data(), sizeof(float4) * tri.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(sc->spos, spos.data(), sizeof(float4) * spos.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(sc->snrm, snrm.data(), sizeof(float4) * snrm.size(), hipMemcpyHostToDevice));
    sc->d.nodes = sc->nodes;
    sc->d.tri = sc->tri;
    sc->d.trest = sc->tri + nt;
    sc->d.spos = sc->spos;
    sc->d.snrm = sc->snrm;
    sc->d.n_nodes = B.n_nodes;
    sc->info.n_nodes = B.n_nodes;
    sc->info.n_triangles = nt;
    sc->info.n_leaves = B.n_leaves;
    sc->info.max_leaf_tris = B.max_leaf;
    sc->info.depth = B.depth;
    sc->info.build_seconds = build_s;
    sc->info.device_bytes = sizeof(float4) * (2 * (size_t)B.n_nodes + tri.size() + spos.size() + snrm.size() +
                                              2 * (size_t)cull.n_nodes) +
                            sizeof(uint32_t) * (parent.size() + cull.ref_leaf.size());
    *out = sc.release();
    return FORNOS_OK;
}

fornos_status fornos_scene_info_get(const fornos_scene* sc, fornos_scene_info* out)
{
    if (!sc || !out) return set_error(FORNOS_INVALID_ARG, "fornos_scene_info_get: null argument");
    *out = sc->info;
    return FORNOS_OK;
}

fornos_status fornos_scene_export_reference_nodes(const fornos_scene* sc, uint32_t* nodes9, uint32_t* ids)
{
    if (!sc) return set_error(FORNOS_INVALID_ARG, "fornos_scene_export_reference_nodes: null scene");
    const FlatBVH& B = sc->bvh;
    if (nodes9) export_reference_nodes(B, nodes9);
    if (ids) std::memcpy(ids, B.soup_tri.data(), sizeof(uint32_t) * B.soup_tri.size());
    return FORNOS_OK;
}

fornos_status fornos_scene_export_soup(const fornos_scene* sc, float* p4, float* n4)
{
    if (!sc) return set_error(FORNOS_INVALID_ARG, "fornos_scene_export_soup: null scene");
    const size_t bytes = sizeof(float4) * 3 * (size_t)sc->info.n_triangles;
    if (p4) HIP_TRY(hipMemcpy(p4, sc->spos, bytes, hipMemcpyDeviceToHost));
    if (n4) HIP_TRY(hipMemcpy(n4, sc->snrm, bytes, hipMemcpyDeviceToHost));
    return FORNOS_OK;
}

void fornos_scene_destroy(fornos_scene* sc) { delete sc; }

fornos_status fornos_meshmap_create(fornos_scene* sc, const fornos_texel_view* tv, int cull,
                                    fornos_meshmap** out)
{
    if (!sc || !tv || !out) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_create: null argument");
    *out = nullptr;
    if (tv->count > 0 && (!tv->positions || !tv->directions))
        return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_create: texel positions/directions missing");
    std::unique_ptr<fornos_meshmap> mm(new fornos_meshmap());
    mm->sc = sc;
    mm->n = tv->count;
    mm->width = tv->width;
    mm->height = tv->height;
    mm->cull = cull ? 1 : 0;
    mm->work_count = div_up(tv->count, kGroup) * kGroup;
    std::vector<float4> pix((size_t)2 * mm->n);
    for (uint32_t i = 0; i < mm->n; ++i) {
        pix[2 * i] = make_float4(tv->positions[3 * i], tv->positions[3 * i + 1], tv->positions[3 * i + 2], 0.0f);
        pix[2 * i + 1] = make_float4(tv->directions[3 * i], tv->directions[3 * i + 1], tv->directions[3 * i + 2], 0.0f);
    }
    if (fornos_status st = dev_alloc(&mm->pix, pix.size())) return st;
    HIP_TRY(hipMemcpy(mm->pix, pix.data(), sizeof(float4) * pix.size(), hipMemcpyHostToDevice));
    if (tv->normals && tv->tangents && tv->bitangents) {
        std::vector<float4> pt((size_t)3 * mm->n);
        for (uint32_t i = 0; i < mm->n; ++i) {
            pt[3 * i] = make_float4(tv->normals[3 * i], tv->normals[3 * i + 1], tv->normals[3 * i + 2], 0.0f);
            pt[3 * i + 1] = make_float4(tv->tangents[3 * i], tv->tangents[3 * i + 1], tv->tangents[3 * i + 2], 0.0f);
            pt[3 * i + 2] = make_float4(tv->bitangents[3 * i], tv->bitangents[3 * i + 1], tv->bitangents[3 * i + 2], 0.0f);
        }
        if (fornos_status st = dev_alloc(&mm->pixt, pt.size())) return st;
        HIP_TRY(hipMemcpy(mm->pixt, pt.data(), sizeof(float4) * pt.size(), hipMemcpyHostToDevice));
    }
    if (fornos_status st = dev_alloc(&mm->coords, mm->n)) return st;
    if (fornos_status st = dev_alloc(&mm->tidx, mm->n)) return st;
    *out = mm.release();
    return FORNOS_OK;
}

static fornos_status meshmap_launch(fornos_meshmap* mm, uint32_t t0, uint32_t t1)
{
    if (t1 <= t0) return FORNOS_OK;
    const uint32_t blocks = div_up(t1 - t0, kBlock);
    if (mm->cull)
        hipLaunchKernelGGL(k_meshmap<true>, dim3(blocks), dim3(kBlock), 0, mm->sc->stream, mm->pix, t0, t1, mm->sc->d, mm->coords, mm->tidx);
    else
        hipLaunchKernelGGL(k_meshmap<false>, dim3(blocks), dim3(kBlock), 0, mm->sc->stream, mm->pix, t0, t1, mm->sc->d, mm->coords, mm->tidx);
    HIP_TRY(hipGetLastError());
    return FORNOS_OK;
}

fornos_status fornos_meshmap_run_step(fornos_meshmap* mm, int* done)
{
    if (!mm) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_run_step: null handle");
    if (mm->offset < mm->work_count) {
        const uint32_t work = std::min(mm->work_count - mm->offset, kWorkPerStep);
        const uint32_t t0 = mm->offset, t1 = std::min(mm->offset + work, mm->n);
        if (fornos_status st = meshmap_launch(mm, t0, t1)) return st;
        mm->offset += work;
    }
    if (done) *done = mm->offset >= mm->work_count;
    return FORNOS_OK;
}

fornos_status fornos_meshmap_run_all(fornos_meshmap* mm, float* ms)
{
    if (!mm) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_run_all: null handle");
    Timer tm(mm->sc->stream);
    if (fornos_status st = tm.start()) return st;
    if (fornos_status st = meshmap_launch(mm, std::min(mm->offset, mm->n), mm->n)) return st;
    mm->offset = mm->work_count;
    return tm.stop(ms);
}

fornos_status fornos_meshmap_reset(fornos_meshmap* mm)
{
    if (!mm) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_reset: null handle");
    mm->offset = 0;
    return FORNOS_OK;
}

fornos_status fornos_meshmap_progress(const fornos_meshmap* mm, float* out)
{
    if (!mm || !out) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_progress: null argument");
    *out = mm->work_count ? (float)mm->offset / (float)mm->work_count : 1.0f;
    return FORNOS_OK;
}

fornos_status fornos_meshmap_read(const fornos_meshmap* mm, float* coords4, uint32_t* tidx)
{
    if (!mm) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_read: null handle");
    HIP_TRY(hipStreamSynchronize(mm->sc->stream));
    if (coords4 && mm->n) HIP_TRY(hipMemcpy(coords4, mm->coords, sizeof(float4) * mm->n, hipMemcpyDeviceToHost));
    if (tidx && mm->n) HIP_TRY(hipMemcpy(tidx, mm->tidx, sizeof(uint32_t) * mm->n, hipMemcpyDeviceToHost));
    return FORNOS_OK;
}

fornos_status fornos_meshmap_texel_count(const fornos_meshmap* mm, uint32_t* out)
{
    if (!mm || !out) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_texel_count: null argument");
    *out = mm->n;
    return FORNOS_OK;
}

void fornos_meshmap_destroy(fornos_meshmap* mm) { delete mm; }

fornos_status fornos_solver_create(int kind, const fornos_solver_params* params, fornos_meshmap* mm,
                                   const float* table, uint32_t n_perm, fornos_solver** out)
{
    if (!params || !mm || !out) return set_error(FORNOS_INVALID_ARG, "fornos_solver_create: null argument");
    *out = nullptr;
    if (kind < FORNOS_SOLVER_HEIGHT || kind > FORNOS_SOLVER_THICKNESS)
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_create: unknown solver kind");
    std::unique_ptr<fornos_solver> s(new fornos_solver());
    s->kind = kind;
    s->p = *params;
    s->mm = mm;
    if ((kind == FORNOS_SOLVER_NORMALS || kind == FORNOS_SOLVER_BENT_NORMALS) && params->tangent_space && !mm->pixt)
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_create: tangent space needs per-texel normals, tangents and bitangents");
    if (s->hemi()) {
        const uint32_t S = params->sample_count;
        if (S == 0 || S > 4096) return set_error(FORNOS_INVALID_ARG, "fornos_solver_create: sample_count must be in [1, 4096]");
        if (!table || n_perm == 0) return set_error(FORNOS_INVALID_ARG, "fornos_solver_create: sample table missing");
        s->P = n_perm;
        s->work_total = (uint64_t)mm->work_count * S;
        std::vector<float4> t4((size_t)n_perm * S);
        for (size_t i = 0; i < t4.size(); ++i) t4[i] = make_float4(table[3 * i], table[3 * i + 1], table[3 * i + 2], 0.0f);
        if (fornos_status st = dev_alloc(&s->samples, t4.size())) return st;
        HIP_TRY(hipMemcpy(s->samples, t4.data(), sizeof(float4) * t4.size(), hipMemcpyHostToDevice));
        if (fornos_status st = dev_alloc(&s->counts, mm->n)) return st;
        if (fornos_status st = dev_alloc(&s->stats, 3)) return st;
        if (fornos_status st = dev_alloc(&s->counter, 1)) return st;
        // per-ray results of one k_rays launch: at most 2^28 rays (256 MiB of
        // hit bytes, 1 GiB of thickness t values)
        s->chunk_texels = std::max<uint32_t>(1, (1u << 28) / S);
        const size_t rays = (size_t)std::min<uint32_t>(s->chunk_texels, std::max<uint32_t>(1, mm->n)) * S;
        const size_t elem = kind == FORNOS_SOLVER_THICKNESS ? sizeof(float) : 1;
        uint8_t* rb = nullptr;
        if (fornos_status st = dev_alloc(&rb, rays * elem)) return st;
        s->raybuf = rb;
    } else {
        s->work_total = mm->work_count;
    }
    if (fornos_status st = dev_alloc(&s->out, (size_t)mm->n * s->comps())) return st;
    *out = s.release();
    return FORNOS_OK;
}

fornos_status fornos_solver_run_step(fornos_solver* s, int* done)
{
    if (!s) return set_error(FORNOS_INVALID_ARG, "fornos_solver_run_step: null handle");
    const uint32_t n = s->mm->n;
    if (s->offset < s->work_total) {
        if (s->hemi()) {
            const uint32_t S = s->p.sample_count;
            // chunks of whole texels, at most 131072 rays (Src/solver_ao.cpp:83-87)
            const uint64_t per = std::max<uint64_t>(1, kWorkPerStep / S) * S;
            const uint64_t work = std::min<uint64_t>(s->work_total - s->offset, per);
            const uint32_t t0 = (uint32_t)(s->offset / S), t1 = (uint32_t)std::min<uint64_t>((s->offset + work) / S, n);
            if (fornos_status st = solver_launch(s, std::min(t0, n), t1, false)) return st;
            s->offset += work;
        } else {
            const uint64_t work = std::min<uint64_t>(s->work_total - s->offset, kWorkPerStep);
            const uint32_t t0 = (uint32_t)s->offset, t1 = (uint32_t)std::min<uint64_t>(s->offset + work, n);
            if (fornos_status st = solver_launch(s, std::min(t0, n), t1, false)) return st;
            s->offset += work;
        }
    }
    if (done) *done = s->offset >= s->work_total;
    return FORNOS_OK;
}

fornos_status fornos_solver_run_all(fornos_solver* s, float* ms)
{
    if (!s) return set_error(FORNOS_INVALID_ARG, "fornos_solver_run_all: null handle");
    const uint32_t n = s->mm->n;
    const uint32_t S = s->hemi() ? s->p.sample_count : 1;
    const uint32_t t0 = (uint32_t)std::min<uint64_t>(s->offset / S, n);
    Timer tm(s->mm->sc->stream);
    if (fornos_status st = tm.start()) return st;
    for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
    s->events.clear();
    s->timed = s->hemi();
    const fornos_status st = solver_launch(s, t0, n, false);
    s->timed = false;
    if (st) return st;
    s->offset = s->work_total;
    if (fornos_status st2 = tm.stop(ms)) return st2;
    s->rays_ms = s->reduce_ms = 0;
    s->launches = (uint32_t)(s->events.size() / 3);
    for (size_t i = 0; i + 2 < s->events.size(); i += 3) {
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, s->events[i], s->events[i + 1]));
        HIP_TRY(hipEventElapsedTime(&b, s->events[i + 1], s->events[i + 2]));
        s->rays_ms += a;
        s->reduce_ms += b;
    }
    return FORNOS_OK;
}

fornos_status fornos_solver_last_timing(const fornos_solver* s, float* rays_ms, float* reduce_ms, uint32_t* launches)
{
    if (!s) return set_error(FORNOS_INVALID_ARG, "fornos_solver_last_timing: null handle");
    if (rays_ms) *rays_ms = s->rays_ms;
    if (reduce_ms) *reduce_ms = s->reduce_ms;
    if (launches) *launches = s->launches;
    return FORNOS_OK;
}

fornos_status fornos_solver_progress(const fornos_solver* s, float* out)
{
    if (!s || !out) return set_error(FORNOS_INVALID_ARG, "fornos_solver_progress: null argument");
    *out = s->work_total ? (float)s->offset / (float)s->work_total : 1.0f;
    return FORNOS_OK;
}

fornos_status fornos_solver_reset(fornos_solver* s)
{
    if (!s) return set_error(FORNOS_INVALID_ARG, "fornos_solver_reset: null handle");
    s->offset = 0;
    return FORNOS_OK;
}

fornos_status fornos_solver_read(const fornos_solver* s, void* out, size_t bytes)
{
    if (!s || !out) return set_error(FORNOS_INVALID_ARG, "fornos_solver_read: null argument");
    const size_t need = sizeof(float) * (size_t)s->mm->n * s->comps();
    if (bytes < need) return set_error(FORNOS_INVALID_ARG, "fornos_solver_read: buffer too small (" + std::to_string(need) + " bytes needed)");
    HIP_TRY(hipStreamSynchronize(s->mm->sc->stream));
    if (need) HIP_TRY(hipMemcpy(out, s->out, need, hipMemcpyDeviceToHost));
    return FORNOS_OK;
}

fornos_status fornos_solver_read_counts(const fornos_solver* s, uint32_t* out)
{
    if (!s || !out) return set_error(FORNOS_INVALID_ARG, "fornos_solver_read_counts: null argument");
    if (!s->hemi()) return set_error(FORNOS_INVALID_ARG, "fornos_solver_read_counts: not a hemisphere solver");
    HIP_TRY(hipStreamSynchronize(s->mm->sc->stream));
    if (s->mm->n) HIP_TRY(hipMemcpy(out, s->counts, sizeof(uint32_t) * s->mm->n, hipMemcpyDeviceToHost));
    return FORNOS_OK;
}

fornos_status fornos_solver_count_work(fornos_solver* s, uint64_t* nodes, uint64_t* tris, uint64_t* rays)
{
    if (!s) return set_error(FORNOS_INVALID_ARG, "fornos_solver_count_work: null handle");
    if (!s->hemi()) return set_error(FORNOS_INVALID_ARG, "fornos_solver_count_work: not a hemisphere solver");
    if (fornos_status st = solver_launch(s, 0, s->mm->n, true)) return st;
    unsigned long long h[3] = {0, 0, 0};
    HIP_TRY(hipStreamSynchronize(s->mm->sc->stream));
    HIP_TRY(hipMemcpy(h, s->stats, sizeof(h), hipMemcpyDeviceToHost));
    if (nodes) *nodes = h[0];
    if (tris) *tris = h[1];
    if (rays) *rays = h[2];
    return FORNOS_OK;
}

void fornos_solver_destroy(fornos_solver* s) { delete s; }

}  // extern "C"
