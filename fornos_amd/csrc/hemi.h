// hemi.h -- hemisphere passes (AO, bent normals, thickness).
//
// k_rays runs persistent waves.  Each lane owns one ray: the ray basis of
// Shaders/ao_step0.comp and one sample of ao/bentnormals/thick_step1.comp.
// A wave owns one texel at a time and hands its samples, in their
// direction-sorted order, to its idle lanes once enough are idle; when they
// run out it takes the next texel of its batch, and a new batch of texels
// from a global counter when that is spent (one atomic per batch of tb
// texels, the pass's hand-out tile: by default one texel at 256 samples and
// more, 256 / S texels below, at most 32).  So a wave never idles
// behind its slowest ray, and its lanes trace rays from few origins, which
// walk the tree in step.  Traversal is "while-while": every lane walks the 4-wide
// culling tree until it holds a leaf (or is done), then the wave tests
// triangles together, so the triangle code never runs for one lane while the
// others wait on nodes.
//
// Exactness.  The culling trees are not the reference's tree, so a hit found
// through them counts only under the reference's rules (chain_ok in
// device.h): AO and BN take a qualifying hit iff its reference leaf is
// reachable; thickness takes a hit at t < best iff its leaf's chain passes
// with min(maxD, t), and otherwise (a box entered no earlier than the hit)
// the ray is redone on the reference tree.  Rays with a direction component
// below 2^-40 use the reference tree and its exact test from the start, and a
// ray whose 4-wide stack would overflow is redone on the stackless binary
// culling tree.  Results therefore equal the reference's ray for ray.
//
// AO walks a tree of its own when the scene gains from it: leaves split into
// groups of at most 2 triangles with their own boxes.  A group whose box
// the ray enters at or beyond maxD holds no hit below maxD, so pruning it is
// exact for AO; BN and thickness count hits beyond maxD in any leaf the
// reference reaches, so they keep whole leaves (whose boxes contain the
// reference's leaf boxes).
//
// Only hits are written: AO adds 1 to its texel's counter, BN sets its
// sample's bit, thickness stores t (the buffer starts at FLT_MAX).  AO and
// BN with the same rays (samples, minD, maxD) can share one traversal
// (kAOBN): both count hits in reachable leaves, AO those at t < maxD and BN
// any, so the AO search, which stops at its first hit, decides both.  k_reduce
// then folds each texel's S results in sample order (step2 + tangentspace),
// the same split as the reference's step1/step2 dispatches.
#pragma once
#include "device.h"

namespace fornos {
namespace {

enum { kAO = 0, kBN = 1, kThick = 2, kAOBN = 3 };  // kAOBN: AO and bent normals from one traversal
enum { kWideTree = 0, kCullTree = 1, kRefTree = 2 };  // traversal modes of a ray
// counting run: box tests, triangle tests, rays, stack overflows, then per wave:
// walk steps, walking lanes, leaf steps, leaf lanes, refills, refilled lanes;
// then thickness rays redone on the reference tree, and rays traced there from
// the start (a direction component below 2^-40)
constexpr uint32_t kStats = 12;
// counting run, per texel: which fallback paths its rays took
enum : uint32_t { kFbOverflow = 1u, kFbRedo = 2u, kFbSlow = 4u };
constexpr uint32_t kRefill = 32;  // default: refill a wave once this many lanes are idle
// Leaf step once the lanes still searching are at most the lanes holding a
// leaf / quorum (0: wait for all).  Thickness lanes that hold a leaf do not
// walk on (its pruning needs the hits first), so they idle while others
// search: its step starts once a third of the lanes still search (knot4096
// thickness 7 % faster than at 8).  Bent normals and the paired pass gain
// 1-2 % at 4; AO is within noise from 4 to 16.
template <int KIND>
constexpr uint32_t quorum()
{
    return KIND == kThick ? 2u : (KIND == kAO ? 8u : 4u);
}
constexpr uint32_t kStack = 16;   // default 4-wide traversal stack entries per lane (16, 24 or 32)

struct RayArgs {
    DevScene sc;
    const float4* __restrict__ coords;
    const uint32_t* __restrict__ tidx;
    const float4* __restrict__ samples;  // per permutation, in ray order; w = sample index
    uint32_t* __restrict__ cnt;         // AO: occluded rays per texel (global texel index)
    uint32_t* __restrict__ mask;        // BN: hit bits, mwords words per texel of the launch
    float* __restrict__ tval;           // thickness: closest t per ray of the launch
    unsigned int* __restrict__ counter;  // next ray of the launch to hand out
    unsigned long long* __restrict__ stats;  // kStats counters (COUNT only)
    uint32_t* __restrict__ fb;          // kFb* flags per texel, global texel index (COUNT only)
    const uint32_t* __restrict__ list;  // null: texels [t0, t0 + ntex); else list[t0 .. t0 + ntex)
    uint32_t S, P, t0, ntex, mwords;    // ntex texels, S rays each
    float mind, maxd;
    uint32_t refill;                    // refill a wave once this many lanes are idle
    uint32_t tb;                        // TB kernels: texels taken from the counter at once (>= 1)
};


// Global index of the launch's q-th texel (the list is wave-uniform, so the
// contiguous case costs one scalar branch).
__device__ inline uint32_t texel_of(const uint32_t* list, uint32_t t0, uint32_t q)
{
    return list ? list[t0 + q] : t0 + q;
}

struct Lane {
    Ray ray;
    uint32_t q, s;      // texel of the launch, sample
    uint32_t cur, sp;   // 4-wide tree: next item (node, kLeafFlag | leaf, kNoChild = pop)
    uint32_t slot;      // 4-wide tree: a found leaf waiting for the wave's leaf step (or kNoChild)
    uint32_t node;      // binary trees: next node
    float best;         // thickness: closest accepted hit
    int mode;
    bool bn;            // kAOBN: a qualifying hit at or beyond maxD was found (bent normals only)
    bool active;
};

template <int KIND>
__device__ inline void finish(const RayArgs& A, Lane& L, bool hit)
{
    if (KIND == kAO) {
        if (hit) atomicAdd(&A.cnt[texel_of(A.list, A.t0, L.q)], 1u);
    } else if (KIND == kBN) {
        if (hit) atomicOr(&A.mask[(size_t)L.q * A.mwords + (L.s >> 5)], 1u << (L.s & 31));
    } else if (KIND == kAOBN) {
        if (hit) atomicAdd(&A.cnt[texel_of(A.list, A.t0, L.q)], 1u);
        if (hit || L.bn) atomicOr(&A.mask[(size_t)L.q * A.mwords + (L.s >> 5)], 1u << (L.s & 31));
    } else {
        if (L.best != FLT_MAX) A.tval[(size_t)L.q * A.S + L.s] = L.best;
    }
    L.active = false;
}

// Triangles [first, first + count) of a reference leaf whose node is `leaf`
// (all of them, or, on AO's tree, those in `mask`: bit j = triangle first +
// j; on the others' tree a non-zero `mask` is the leaf's group node + 1), under the
// hit rules above.  Returns 1 when the ray is decided (AO/BN hit), 2 when it
// must be redone on the reference tree, 0 otherwise.  kAOBN: a hit in a
// reachable leaf at t >= maxD decides bent normals only (L.bn) and the search
// for an AO hit (t < maxD) goes on.
template <int KIND, bool COUNT>
__device__ inline int leaf_hits(const RayArgs& A, Lane& L, uint32_t first, uint32_t count, uint32_t leaf,
                                uint32_t mask, uint32_t& ct)
{
    const uint32_t end = first + count;
    const float tri_hi = KIND == kAO ? A.maxd : INFINITY;
    bool reach = false;  // kAOBN: the leaf's chain is known to pass
    // a hit at t (FLT_MAX: none) -> -1 to go on, or the result
    auto take = [&](float t) -> int {
        if (t == FLT_MAX) return -1;
        if (KIND == kAOBN) {
            if (!reach) {
                if (!chain_ok(L.ray, A.sc, leaf, A.maxd)) return 0;  // one chain per leaf
                reach = true;
            }
            if (t < A.maxd) return 1;
            L.bn = true;
            return -1;
        }
        if (KIND != kThick) return chain_ok(L.ray, A.sc, leaf, A.maxd) ? 1 : 0;  // one chain per leaf
        if (chain_ok(L.ray, A.sc, leaf, fminf(A.maxd, t))) {
            L.best = t;
            return -1;
        }
        return chain_ok(L.ray, A.sc, leaf, A.maxd) ? 2 : 0;
    };
    if (KIND != kAO && mask) {
        // group boxes inside the whole leaf: a group whose padded box the ray
        // enters no earlier than the pass's bound holds no hit that counts
        // (thickness: its best hit; BN: hits at any distance count, so only
        // groups the ray misses are skipped)
        const float4* nd = A.sc.wsub + 8 * (size_t)(mask - 1);
        const float4 lx = nd[0], ly = nd[1], lz = nd[2], hx = nd[3], hy = nd[4], hz = nd[5];
        const uint4 gm = *(const uint4*)(nd + 6);
        const float T = KIND == kThick ? L.best : INFINITY;
        const float4 lo[4] = {make_float4(lx.x, ly.x, lz.x, 0), make_float4(lx.y, ly.y, lz.y, 0),
                              make_float4(lx.z, ly.z, lz.z, 0), make_float4(lx.w, ly.w, lz.w, 0)};
        const float4 hi[4] = {make_float4(hx.x, hy.x, hz.x, 0), make_float4(hx.y, hy.y, hz.y, 0),
                              make_float4(hx.z, hy.z, hz.z, 0), make_float4(hx.w, hy.w, hz.w, 0)};
        const uint32_t gms[4] = {gm.x, gm.y, gm.z, gm.w};
        mask = 0;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            float al;
            if ((gms[g] != 0) & box_maybe(L.ray, lo[g], hi[g], T, al)) mask |= gms[g];
        }
        if (!mask) return 0;
    }
    if (mask) {
        while (mask) {
            const uint32_t k = first + (uint32_t)__builtin_ctz(mask);
            mask &= mask - 1;
            if (COUNT) ++ct;
            const int r = take(hemi_tri(L.ray.o, L.ray.d, k, A.sc, A.mind, KIND == kThick ? L.best : tri_hi));
            if (r >= 0) return r;
        }
        return 0;
    }
    for (uint32_t k0 = first; k0 < end; k0 += 4) {
        // the leaf's (n, a.n) records are contiguous: fetch four at once
        float4 q[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) q[j] = k0 + j < end ? A.sc.tri[k0 + j] : make_float4(0, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t k = k0 + j;
            if (k >= end) continue;
            if (COUNT) ++ct;
            const int r = take(hemi_tri_q(L.ray.o, L.ray.d, q[j], k, A.sc, A.mind, KIND == kThick ? L.best : tri_hi));
            if (r >= 0) return r;
        }
    }
    return 0;
}

// One node of the 4-wide culling tree (or one stack pop); STK stack entries.
template <int KIND, bool COUNT, int STK>
__device__ inline void wide_step(const RayArgs& A, Lane& L, uint32_t* stk_node, uint32_t& cn, uint32_t& nov)
{
    const uint32_t tid = threadIdx.x;
    const float T = KIND == kThick ? fminf(L.best, A.maxd) : A.maxd;
    if (L.cur == kNoChild) {
        // pop, and visit the popped node in the same step (a popped leaf
        // ends the walk).  Thickness keeps no entry distances on the stack:
        // a subtree entered beyond its best hit is visited and all of its
        // children fail the test against that bound, which costs less than
        // the occupancy a second LDS stack takes.
        if (L.sp == 0) {
            finish<KIND>(A, L, false);
            return;
        }
        --L.sp;
        L.cur = stk_node[L.sp * kBlock + tid];
        if (L.cur == kNoChild || (L.cur & kLeafFlag)) return;
    }
    const float4* nd = (KIND == kAO ? A.sc.wide_ao : A.sc.wide) + 8 * (size_t)L.cur;
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
        if (COUNT && chs[k] != kNoChild) ++cn;
        float al = 0;
        // `&`, not `&&`: the four tests run without a branch each (an empty
        // slot's box is in the node and harmless to test)
        const bool pass = (chs[k] != kNoChild) & box_maybe(L.ray, lo[k], hi[k], T, al);
        e[k] = pass ? al : INFINITY;
        c[k] = pass ? chs[k] : kNoChild;
    }
    // nearest first: sort the four (entry, child) pairs
#define FORNOS_CSWAP(i, j)                                            \
    if (e[j] < e[i]) {                                                \
        const float te = e[i]; e[i] = e[j]; e[j] = te;                \
        const uint32_t tc = c[i]; c[i] = c[j]; c[j] = tc;             \
    }
    FORNOS_CSWAP(0, 1) FORNOS_CSWAP(2, 3) FORNOS_CSWAP(0, 2) FORNOS_CSWAP(1, 3) FORNOS_CSWAP(1, 2)
#undef FORNOS_CSWAP
    L.cur = c[0];
    // the sort put the passing children first (a failed one sorts last at
    // INFINITY): c[1..n] go on the stack, the nearest on top
    const uint32_t n = (uint32_t)(c[1] != kNoChild) + (uint32_t)(c[2] != kNoChild) + (uint32_t)(c[3] != kNoChild);
    if (L.sp + n > STK) {  // redo the ray on the stackless binary culling tree
        if (COUNT) {
            ++nov;
            atomicOr(&A.fb[texel_of(A.list, A.t0, L.q)], kFbOverflow);
        }
        L.slot = kNoChild;
        L.mode = kCullTree;
        L.node = 0;
        L.best = FLT_MAX;
        return;
    }
#pragma unroll
    for (uint32_t k = 1; k <= 3; ++k)
        if (k <= n) stk_node[(L.sp + n - k) * kBlock + tid] = c[k];
    L.sp += n;
}

// One node of a binary tree in skip-link order: the culling tree (boxes
// tested conservatively, hit rules above) or the reference tree (the
// reference's own test and traversal, Shaders/ao_step1.comp:166-193).
template <int KIND, bool COUNT>
__device__ inline void binary_step(const RayArgs& A, Lane& L, uint32_t& cn, uint32_t& ct, uint32_t& nref)
{
    const bool ref = L.mode == kRefTree;
    const float4* nodes = ref ? A.sc.nodes : A.sc.cnodes;
    const uint32_t n_nodes = ref ? A.sc.n_nodes : A.sc.n_cnodes;
    const float4 lo = nodes[2 * L.node], hi = nodes[2 * L.node + 1];
    const uint32_t w = __float_as_uint(hi.w);
    const float T = KIND == kThick ? fminf(L.best, A.maxd) : A.maxd;
    if (COUNT) ++cn;
    const bool visit = ref ? node_visit(L.ray, lo, hi, T) : node_maybe(L.ray, lo, hi, T);
    if (visit) {
        if (w & kLeafFlag) {
            const uint32_t first = __float_as_uint(lo.w), end = first + (w & ~kLeafFlag);
            if (ref) {
                const float tri_hi = KIND == kAO ? A.maxd : (KIND == kThick ? L.best : INFINITY);
                for (uint32_t k = first; k < end; ++k) {
                    if (COUNT) ++ct;
                    const float t = hemi_tri(L.ray.o, L.ray.d, k, A.sc, A.mind, KIND == kThick ? L.best : tri_hi);
                    if (t == FLT_MAX) continue;
                    if (KIND == kAOBN && !(t < A.maxd)) {
                        L.bn = true;
                        continue;
                    }
                    if (KIND != kThick) {
                        finish<KIND>(A, L, true);
                        return;
                    }
                    L.best = t;
                }
            } else {
                const int r = leaf_hits<KIND, COUNT>(A, L, first, w & ~kLeafFlag, A.sc.cref[L.node], 0u, ct);
                if (r == 1) {
                    finish<KIND>(A, L, true);
                    return;
                }
                if (r == 2) {
                    if (COUNT) {
                        ++nref;
                        atomicOr(&A.fb[texel_of(A.list, A.t0, L.q)], kFbRedo);
                    }
                    L.mode = kRefTree;
                    L.node = 0;
                    L.best = FLT_MAX;
                    return;
                }
            }
        }
        ++L.node;
    } else {
        L.node = (w & kLeafFlag) ? L.node + 1 : w;
    }
    if (L.node >= n_nodes) finish<KIND>(A, L, false);
}

// BATCH: `jobs` holds njobs independent passes (each its own scene, texels,
// samples, counter and outputs).  A wave works on one job at a time -- the
// job index is wave-uniform, so its parameters live in scalar registers --
// and moves on once the job's counter is exhausted and its lanes are idle.
// Waves per SIMD the kernel is compiled for: as many as the LDS stack leaves
// room for (160 KiB per CU; 8 at 16 entries, 6 at 24, 5 at 32).  At 8 the
// compiler is held to 64 VGPRs (AO needs no spill at that budget; BN,
// thickness and the paired kind spill 16-36 bytes per lane) and the latency
// of divergent node and triangle fetches is hidden by more waves -- 8-16 %
// faster than the 6-7 waves the unconstrained register allocation leaves
// room for.
// A job's arguments, word by word through readfirstlane: every wave reads
// the same job, and saying so keeps them in scalar registers.
__device__ inline RayArgs load_job(const RayArgs* jobs, uint32_t j)
{
    static_assert(sizeof(RayArgs) % 4 == 0, "RayArgs is whole words");
    RayArgs A;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(jobs + j);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&A);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(RayArgs) / 4; ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return A;
}

template <int STK>
constexpr int waves_for_stack()
{
    constexpr uint32_t blocks = (160u * 1024u) / (STK * kBlock * 4u);  // per CU, by LDS
    constexpr uint32_t waves = blocks * (kBlock / 64u) / 4u;            // per SIMD (4 per CU)
    return waves > 8 ? 8 : (int)waves;
}

// TB: texels are taken from the launch's counter A.tb at a time (the pass's
// hand-out tile, > 1), with the wave's texel state in LDS; otherwise one
// texel per atomic with the state in scalar registers, which leaves the
// register allocation of the walk as it is (the paired AO + BN pass at 256
// samples ran 13 % slower with the batched hand-out's state).
template <int KIND, bool COUNT, bool BATCH, int STK, bool TB>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for_stack<STK>()))) void k_rays(RayArgs A0,
                                                                                         const RayArgs* __restrict__ jobs,
                                                                                         uint32_t njobs)
{
    __shared__ uint32_t stk_node[STK * kBlock];
    uint32_t job = 0;
    RayArgs A = BATCH ? load_job(jobs, 0) : A0;
    // per wave: its texel (launch index) wq and the next of its samples to
    // hand out ws (ws >= S: none left); TB: also its batch of texels from the
    // launch's counter, [bq, bend) not yet taken.  TB keeps the state in LDS,
    // as it is used only at refills (the batch kernel keeps it in scalar
    // registers: in LDS its job arguments left them and spilled).
    constexpr bool kLdsState = TB && !BATCH;
    uint32_t wq = 0, ws = A.S, bq = 0, bend = 0;
    __shared__ uint32_t wstate[kBlock / 64][4];
    // TB thickness: the mapping results (hit triangle or kMiss) of the
    // wave's batch, loaded by its lanes at once when the batch is taken, at
    // slot (texel & 31) (a batch is at most 32 consecutive texels), so taking
    // a texel reads LDS rather than waiting on a load of its own (thickness-16
    // on heightfield2048 15.9 -> 14.2 ms).  AO, bent normals and the paired
    // pass keep the load per texel: with the array they ran 0.5-1 % slower.
    constexpr bool kBatchLds = TB && KIND == kThick;
    __shared__ uint32_t wbatch[kBatchLds ? kBlock / 64 : 1][32];
    const uint32_t lane = threadIdx.x & 63;
    uint32_t* const ws_of = wstate[threadIdx.x >> 6];
    uint32_t* const bt_of = wbatch[kBatchLds ? threadIdx.x >> 6 : 0];
    if (kLdsState && lane == 0) {
        ws_of[0] = 0;
        ws_of[1] = 0xffffffffu;
        ws_of[2] = 0;
        ws_of[3] = 0;
    }
    Lane L;
    L.ray.o = L.ray.d = L.ray.inv = L.ray.oi = mk(0, 0, 0);
    L.ray.slow = false;
    L.q = L.s = L.cur = L.sp = L.node = 0;
    L.slot = kNoChild;
    L.best = FLT_MAX;
    L.mode = kWideTree;
    L.active = false;
    L.bn = false;
    bool exhausted = false;
    uint32_t cn = 0, ct = 0, nr = 0, nov = 0, nref = 0, nslow = 0;
    unsigned long long w_steps = 0, w_lanes = 0, l_steps = 0, l_lanes = 0, r_steps = 0, r_lanes = 0;
    for (;;) {
        // refill idle lanes from the wave's texel: a wave hands its lanes
        // the samples of one texel (rays from one origin, in the samples'
        // direction-sorted order) and takes the next texel of its batch when
        // they run out, and the next batch from the launch's counter (one
        // atomic per batch) when the batch is spent
        const uint64_t idle = __ballot(!L.active);
        const uint32_t nidle = __popcll(idle);
        if (!exhausted && (nidle >= A.refill || nidle == 64)) {
            if (COUNT && lane == 0) {
                ++r_steps;
                r_lanes += nidle;
            }
            const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1));
            if (kLdsState) {
                wq = __builtin_amdgcn_readfirstlane(ws_of[0]);
                ws = __builtin_amdgcn_readfirstlane(ws_of[1]);
                bq = __builtin_amdgcn_readfirstlane(ws_of[2]);
                bend = __builtin_amdgcn_readfirstlane(ws_of[3]);
            }
            uint32_t my_q = kMiss, my_s = 0;
            uint32_t filled = 0;
            while (filled < nidle) {
                if (ws >= A.S && !TB) {
                    uint32_t nq = 0;
                    if (lane == 0) nq = atomicAdd(A.counter, 1u);
                    nq = __builtin_amdgcn_readfirstlane(nq);  // wave-uniform, so the texel state stays scalar
                    if (nq >= A.ntex) {
                        exhausted = true;
                        break;
                    }
                    wq = nq;
                    // a texel whose mapping missed casts no rays (the load is
                    // the same for every lane: readfirstlane keeps the state scalar)
                    ws = __builtin_amdgcn_readfirstlane(A.tidx[texel_of(A.list, A.t0, wq)]) == kMiss ? A.S : 0;
                    continue;
                }
                if (ws >= A.S) {
                    if (bq >= bend) {
                        // a tile of 0 would never advance the counter (the host
                        // keeps tiles at most 32, which the batch's LDS slots need)
                        const uint32_t tb = max(A.tb, 1u);
                        uint32_t nq = 0;
                        if (lane == 0) nq = atomicAdd(A.counter, tb);
                        nq = __builtin_amdgcn_readfirstlane(nq);  // wave-uniform, so the texel state stays scalar
                        if (nq >= A.ntex) {
                            exhausted = true;
                            break;
                        }
                        bq = nq;
                        bend = min(nq + tb, A.ntex);
                        if (kBatchLds && lane < bend - bq)
                            bt_of[(bq + lane) & 31] = A.tidx[texel_of(A.list, A.t0, bq + lane)];
                    }
                    wq = bq++;
                    const uint32_t t = kBatchLds ? bt_of[wq & 31] : A.tidx[texel_of(A.list, A.t0, wq)];
                    ws = __builtin_amdgcn_readfirstlane(t) == kMiss ? A.S : 0;
                    continue;
                }
                const uint32_t take = min(nidle - filled, A.S - ws);
                if (!L.active && rank >= filled && rank < filled + take) {
                    my_q = wq;
                    my_s = ws + (rank - filled);
                }
                filled += take;
                ws += take;
            }
            if (kLdsState && lane == 0) {
                ws_of[0] = wq;
                ws_of[1] = ws;
                ws_of[2] = bq;
                ws_of[3] = bend;
            }
            if (!L.active && my_q != kMiss) {
                const uint32_t q = my_q, s = my_s, g = texel_of(A.list, A.t0, q);
                const uint32_t t = A.tidx[g];
                const Basis B = ray_basis(A.sc, A.coords[g], t);
                const V3 d = KIND == kThick ? -B.d : B.d;
                L.ray.o = B.o;
                const float4 rs = A.samples[(size_t)(g % A.P) * A.S + s];
                L.ray.d = sample_dir(B, d, rs);
                L.ray.inv = mk(1.0f / L.ray.d.x, 1.0f / L.ray.d.y, 1.0f / L.ray.d.z);
                L.ray.oi = mk(L.ray.o.x * L.ray.inv.x, L.ray.o.y * L.ray.inv.y, L.ray.o.z * L.ray.inv.z);
                L.ray.slow = !(fabsf(L.ray.d.x) >= kSlowDir && fabsf(L.ray.d.y) >= kSlowDir &&
                               fabsf(L.ray.d.z) >= kSlowDir);
                L.mode = L.ray.slow ? kRefTree : kWideTree;
                if (COUNT && L.ray.slow) {
                    ++nslow;
                    atomicOr(&A.fb[g], kFbSlow);
                }
                L.q = q;
                L.s = __float_as_uint(rs.w);  // the sample's index in the table
                L.cur = 0;
                L.sp = 0;
                L.slot = kNoChild;
                L.node = 0;
                L.best = FLT_MAX;
                L.bn = false;
                L.active = true;
                if (COUNT) ++nr;
            }
        }
        if (exhausted && __ballot(L.active) == 0) {
            if (!BATCH) break;
            job = __builtin_amdgcn_readfirstlane(job + 1);
            if (job >= njobs) break;
            A = load_job(jobs, job);
            ws = A.S;
            bq = bend = 0;
            if (kLdsState && lane == 0) {
                ws_of[1] = 0xffffffffu;
                ws_of[3] = 0;
            }
            exhausted = false;
            continue;
        }
        // phase 1: walk the 4-wide tree until every lane has parked a leaf (or
        // is done).  Speculative: a lane that has parked one keeps walking
        // while others search, and stops at its next leaf.
        for (;;) {
            const bool wide = L.active && L.mode == kWideTree;
            const bool at_leaf = L.cur != kNoChild && (L.cur & kLeafFlag);
            if (wide && at_leaf && L.slot == kNoChild) {
                L.slot = L.cur;
                L.cur = kNoChild;
            }
            const uint64_t need = __ballot(wide && L.slot == kNoChild);
            if (need == 0) break;
            if (quorum<KIND>()) {
                // leaf step once few lanes are still searching (they go on
                // walking after it); never without a parked leaf
                const uint64_t parked = __ballot(wide && L.slot != kNoChild);
                if (parked != 0 && __popcll(need) * quorum<KIND>() <= (uint32_t)__popcll(parked)) break;
            }
            // thickness does not walk past its leaf: its pruning needs the hits first
            const bool walk = wide && (KIND != kThick || L.slot == kNoChild) &&
                              !(L.cur != kNoChild && (L.cur & kLeafFlag)) &&
                              !(L.cur == kNoChild && L.sp == 0 && L.slot != kNoChild);
            if (COUNT) {
                const uint64_t wb = __ballot(walk);
                if (lane == 0) {
                    ++w_steps;
                    w_lanes += __popcll(wb);
                }
            }
            if (walk) wide_step<KIND, COUNT, STK>(A, L, stk_node, cn, nov);
        }
        // phase 2: parked leaves together; rays on a binary tree take one step
        if (COUNT) {
            const uint64_t lb = __ballot(L.active);
            if (lane == 0) {
                ++l_steps;
                l_lanes += __popcll(lb);
            }
        }
        if (L.active) {
            if (L.mode != kWideTree) {
                binary_step<KIND, COUNT>(A, L, cn, ct, nref);
            } else if (L.slot != kNoChild) {  // (a lane still searching walks on)
                const uint4 lf = (KIND == kAO ? A.sc.wleaf_ao : A.sc.wleaf)[L.slot & ~kLeafFlag];
                L.slot = kNoChild;
                const int r = leaf_hits<KIND, COUNT>(A, L, lf.x, lf.y, lf.z, lf.w, ct);
                if (r == 1) {
                    finish<KIND>(A, L, true);
                } else if (r == 2) {
                    if (COUNT) {
                        ++nref;
                        atomicOr(&A.fb[texel_of(A.list, A.t0, L.q)], kFbRedo);
                    }
                    L.mode = kRefTree;
                    L.node = 0;
                    L.best = FLT_MAX;
                }
            }
        }
    }
    if (COUNT) {
        unsigned long long a = cn, b = ct, c = nr, o = nov, f = nref, sl = nslow;
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_down(a, off);
            b += __shfl_down(b, off);
            c += __shfl_down(c, off);
            o += __shfl_down(o, off);
            f += __shfl_down(f, off);
            sl += __shfl_down(sl, off);
        }
        if (lane == 0) {
            atomicAdd(&A.stats[4], w_steps);
            atomicAdd(&A.stats[5], w_lanes);
            atomicAdd(&A.stats[6], l_steps);
            atomicAdd(&A.stats[7], l_lanes);
            atomicAdd(&A.stats[8], r_steps);
            atomicAdd(&A.stats[9], r_lanes);
            atomicAdd(&A.stats[3], o);
            atomicAdd(&A.stats[10], f);
            atomicAdd(&A.stats[11], sl);
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
    const uint32_t* __restrict__ mask;
    const float* __restrict__ tval;
    float* __restrict__ out;            // 1 or 3 floats per texel (global texel index)
    uint32_t* __restrict__ counts;      // per texel (AO: already counted by k_rays)
    const uint32_t* __restrict__ list;  // as RayArgs::list
    uint32_t S, P, t0, ntex, mwords;
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
    const uint32_t g = texel_of(A.list, A.t0, q), S = A.S;
    const uint32_t t = A.tidx[g];
    if (KIND == kAO) {
        // the reference's sequential float sum of 0/1 values is the exact count
        A.out[g] = t == kMiss ? 1.0f : 1.0f - (float)A.counts[g] / (float)S;
        return;
    }
    uint32_t cnt = 0;
    if (KIND == kThick) {
        float acc = 0;
        if (t != kMiss) {
            const float* tv = A.tval + (size_t)q * S;
            // unrolled so several loads are in flight (the sums keep sample order)
#pragma unroll 8
            for (uint32_t s = 0; s < S; ++s) {
                const float v = tv[s];
                if (v != FLT_MAX) ++cnt;
                acc += v != FLT_MAX ? v : A.maxd;
            }
        }
        A.out[g] = t == kMiss ? A.maxd : acc / (float)S;
    } else {
        V3 r = mk(0, 0, 0);
        if (t != kMiss) {
            const Basis B = ray_basis(A.sc, A.coords[g], t);
            const float4* smp = A.samples + (size_t)(g % A.P) * S;
            const uint32_t* m = A.mask + (size_t)q * A.mwords;
            V3 acc = mk(0, 0, 0);
#pragma unroll 8
            for (uint32_t s = 0; s < S; ++s) {
                if ((m[s >> 5] >> (s & 31)) & 1u) {  // adds vec3(0): acc never holds -0
                    ++cnt;
                    continue;
                }
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

// Fills n 32-bit words with v (a kernel rather than hipMemsetAsync, so the
// whole pass can be captured into a HIP graph).
__global__ void k_fill(uint32_t* p, size_t n, uint32_t v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Sort key of texel g for the order its rays are handed out in: the Morton
// code of its hit position, 10 bits per axis over the scene's box (the
// reference tree's root); texels whose mapping missed (no rays) sort last.
__device__ inline uint32_t spread3(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_order_keys(DevScene sc, const float4* __restrict__ coords, const uint32_t* __restrict__ tidx,
                             uint32_t n, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    vals[g] = g;
    const uint32_t t = tidx[g];
    if (t == kMiss || sc.n_nodes == 0) {
        keys[g] = 0xffffffffu;
        return;
    }
    const float4 c = coords[g];
    const float4 a = sc.spos[t], b = sc.spos[t + 1], d = sc.spos[t + 2];
    const float p[3] = {a.x * c.y + b.x * c.z + d.x * c.w, a.y * c.y + b.y * c.z + d.y * c.w,
                        a.z * c.y + b.z * c.z + d.z * c.w};
    const float4 lo = sc.nodes[0], hi = sc.nodes[1];
    const float l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
    uint32_t key = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float u = fminf(fmaxf((p[k] - l[k]) / fmaxf(h[k] - l[k], 1e-30f), 0.0f), 1.0f);  // NaN -> 0
        key |= spread3((uint32_t)(u * 1023.0f)) << k;
    }
    keys[g] = key;
}

// p[list[i]] = v for the n listed words.
__global__ void k_fill_list(uint32_t* p, const uint32_t* list, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[list[i]] = v;
}

}  // namespace
}  // namespace fornos
