// kernels.hip -- gfx950 kernels of the baking hot path and the device side of
// the C ABI (include/fornos_hip.h).
//
// The kernels live in device.h (layout, predicates), mapping.h, hemi.h and
// texel.h; this file holds the handles and the launches.
//
// Everything is compiled with -ffp-contract=off and IEEE-correct fp32
// division/sqrt so that results are bit-identical to the CPU restatement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "device.h"
#include "hemi.h"
#include "mapping.h"
#include "texel.h"
#include "raster.h"
#include "export.h"
#include "hiphost.h"
#include "../../include/fornos_denoise.h"
#include "../../include/fornos_export.h"

#include <hipcub/hipcub.hpp>
#include <mutex>

#pragma clang fp contract(off)

using namespace fornos;



// AO's split-leaf tree: per-leaf split threshold (triangle tests), build it
// when the predicted cost is below kSplitPredict of the plain tree's, keep it
// when its own cost is below kSplitKeep of it.
constexpr float kSplitCost = 1.0f;
constexpr double kSplitPredict = 0.95, kSplitKeep = 0.9;

// Shared ownership across the C ABI, as the reference's shared_ptr members (a
// solver holds its MeshMapping, Src/solver_ao.h:93): an object made from
// another holds a reference to it, *_destroy drops the caller's, and the last
// reference frees it.  So handles may be destroyed in any order -- a garbage
// collector may finalize a scene before the mapping that uses it.
template <class T>
void retain(T* p)
{
    if (p) p->refs.fetch_add(1, std::memory_order_relaxed);
}
template <class T>
void release(T* p)
{
    if (p && p->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) delete p;
}

struct fornos_scene {
    std::atomic<int> refs{1};
    int device = 0;
    uint32_t cus = 256;  // compute units of the scene's device
    hipStream_t stream = nullptr;
    DevScene d{};
    float4 *nodes = nullptr, *tri = nullptr, *spos = nullptr, *snrm = nullptr, *cnodes = nullptr;
    uint32_t *chain = nullptr, *cref = nullptr;
    float4* wide = nullptr;
    uint4* wleaf = nullptr;
    float4* wide_ao = nullptr;  // AO's tree with split leaves (null: AO walks `wide`)
    uint4* wleaf_ao = nullptr;
    float4* wsub = nullptr;     // group nodes inside the whole leaves of `wide`
    FlatBVH bvh;
    fornos_scene_info info{};
    // image export: its own stream (ordered after this scene's work when an
    // export starts, so the next pass can run meanwhile) and a device
    // workspace kept between exports; one export at a time
    std::mutex xmu;
    hipStream_t xstream = nullptr;
    hipEvent_t xevent = nullptr;
    uint8_t* xbuf = nullptr;
    size_t xcap = 0;
    ~fornos_scene()
    {
        if (xstream) (void)hipStreamSynchronize(xstream);
        if (stream) (void)hipStreamSynchronize(stream);
        dev_free(xbuf);
        if (xevent) (void)hipEventDestroy(xevent);
        if (xstream) (void)hipStreamDestroy(xstream);
        dev_free(nodes); dev_free(tri); dev_free(spos); dev_free(snrm);
        dev_free(cnodes); dev_free(chain); dev_free(cref); dev_free(wide); dev_free(wleaf); dev_free(wide_ao); dev_free(wleaf_ao); dev_free(wsub);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct fornos_meshmap {
    std::atomic<int> refs{1};
    fornos_scene* sc = nullptr;  // one reference held
    uint32_t n = 0, width = 0, height = 0;
    uint32_t work_count = 0;  // n rounded up to 64, Src/meshmapping.cpp:165
    uint32_t offset = 0;      // texels mapped so far
    int cull = 1;
    int predicate = FORNOS_PREDICATE_REFERENCE;  // triangle test of the mapping
    int traversal = FORNOS_TRAVERSAL_WIDE;        // tree the reference predicate's passes walk
    uint32_t wide_stack = kMapStack;              // stack entries the wide passes may use
    uint32_t* redo = nullptr;  // passes redone on the reference tree: undecidable hit, stack overflow
    Marks marks;              // parts of the last fornos_meshmap_submit
    uint32_t mark_from = 0;   // texel its first part started at
    float4 *pix = nullptr, *pixt = nullptr, *coords = nullptr;
    uint32_t *tidx = nullptr, *indices = nullptr;
    uint32_t* order = nullptr;  // texels by the Morton code of their hit positions
    bool order_valid = false;   // built from the current mapping results
    ~fornos_meshmap()
    {
        if (sc && sc->stream) (void)hipStreamSynchronize(sc->stream);
        dev_free(pix); dev_free(pixt); dev_free(coords); dev_free(tidx); dev_free(indices); dev_free(order);
        dev_free(redo);
        release(sc);
    }
};

struct fornos_solver {
    std::atomic<int> refs{1};
    int kind = 0;
    fornos_solver_params p{};
    fornos_meshmap* mm = nullptr;  // one reference held
    uint32_t P = 0;
    uint64_t work_total = 0, offset = 0;  // work items (rays or texels)
    float4* samples = nullptr;      // sample table, reference order (reductions)
    float4* ray_samples = nullptr;  // the same samples sorted by direction, w = sample index
    float* out = nullptr;
    uint32_t* counts = nullptr;
    unsigned long long* stats = nullptr;  // kStats counters of the counting run
    uint32_t* fbflags = nullptr;          // per-texel fallback flags of the counting run
    uint32_t* dlist = nullptr;            // texel list of the last fornos_solver_run_texels
    uint32_t dlist_cap = 0;
    void* raybuf = nullptr;            // per-ray results of one launch (hit bytes or t)
    uint32_t chunk_texels = 0;         // texels per k_rays launch
    uint32_t blocks_per_cu = 8, refill = kRefill, stack = kStack;  // k_rays launch shape
    uint32_t tile = 0;                      // texels per hand-out (0: default_tile)
    int ray_order = FORNOS_RAY_ORDER_AUTO;  // order the texels' rays are handed out in
    uint32_t* counter = nullptr;       // k_rays work counter
    // per-launch events of the last timed run: (before k_rays, after k_rays, after k_reduce)
    std::vector<hipEvent_t> events;
    bool timed = false;
    Marks marks;              // parts of the last fornos_solver_submit
    uint32_t mark_from = 0;   // texel its first part started at
    uint64_t table_hash = 0;  // FNV-1a of the sample table (hemisphere solvers)
    fornos_solver* partner = nullptr;  // AO: the BN solver its passes also compute
    fornos_solver* owner = nullptr;    // BN: the AO solver it is paired with
    hipEvent_t done_ev = nullptr;      // recorded after the solver's last queued work (exports wait on it)
    float rays_ms = 0, reduce_ms = 0;
    uint32_t launches = 0;
    ~fornos_solver()
    {
        if (partner) partner->owner = nullptr;
        if (owner) owner->partner = nullptr;
        if (mm && mm->sc && mm->sc->stream) (void)hipStreamSynchronize(mm->sc->stream);
        dev_free(samples); dev_free(ray_samples); dev_free(out); dev_free(counts); dev_free(stats);
        dev_free(fbflags); dev_free(dlist);
        dev_free(raybuf); dev_free(counter);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (done_ev) (void)hipEventDestroy(done_ev);
        release(mm);
    }
    bool hemi() const { return kind >= FORNOS_SOLVER_AO; }
    uint32_t comps() const
    {
        return (kind == FORNOS_SOLVER_POSITION || kind == FORNOS_SOLVER_NORMALS || kind == FORNOS_SOLVER_BENT_NORMALS) ? 3 : 1;
    }
};

// Scratch of fornos_solver_run_all_noalloc, the entry point that allocates
// nothing: the k_rays work counter in the first kScratchHead bytes, then one
// launch's bent-normal hit bits or thickness distances.  Each device holds its
// own copy at its own address for the life of the process; host code reaches
// it only through the address hipGetSymbolAddress returns for that device
// (scratch_for), never through the array's name.
constexpr size_t kScratchBytes = 64ull << 20, kScratchHead = 256, kScratchData = kScratchBytes - kScratchHead;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch[kScratchDevices] = {};

// The device address of `scratch` on `device`, looked up once per device
// with that device current.
fornos_status scratch_for(int device, uint8_t** out)
{
    if (device < 0 || device >= kScratchDevices)
        return set_error(FORNOS_UNSUPPORTED, "fornos_solver_run_all_noalloc: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        int cur = 0;
        HIP_TRY(hipGetDevice(&cur));
        HIP_TRY(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        const hipError_t e2 = hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return set_error(FORNOS_HIP_ERROR, std::string("fornos_solver_run_all_noalloc: hipGetSymbolAddress(scratch): ") +
                                                   hipGetErrorString(e));
        HIP_TRY(e2);
        g_scratch[device] = (uint8_t*)p;
    }
    *out = g_scratch[device];
    return FORNOS_OK;
}

// Launch a hemisphere kernel over texels [t0, t1).
// k_rays with a stack of 16, 24 or 32 entries per lane; tb: texels handed out
// in batches (passes of fewer than 256 samples; not in the counting run)
template <int KIND, bool COUNT, bool BATCH, bool TB>
void launch_rays_tb(uint32_t stack, uint32_t blocks, hipStream_t st, const RayArgs& A, const RayArgs* jobs, uint32_t nj)
{
    if (stack == 16)
        hipLaunchKernelGGL((k_rays<KIND, COUNT, BATCH, 16, TB>), dim3(blocks), dim3(kBlock), 0, st, A, jobs, nj);
    else if (stack == 32)
        hipLaunchKernelGGL((k_rays<KIND, COUNT, BATCH, 32, TB>), dim3(blocks), dim3(kBlock), 0, st, A, jobs, nj);
    else
        hipLaunchKernelGGL((k_rays<KIND, COUNT, BATCH, 24, TB>), dim3(blocks), dim3(kBlock), 0, st, A, jobs, nj);
}

template <int KIND, bool COUNT, bool BATCH>
void launch_rays(uint32_t stack, uint32_t blocks, hipStream_t st, const RayArgs& A, const RayArgs* jobs, uint32_t nj,
                 bool tb = false)
{
    if (tb && !COUNT)
        launch_rays_tb<KIND, false, BATCH, true>(stack, blocks, st, A, jobs, nj);
    else
        launch_rays_tb<KIND, COUNT, BATCH, false>(stack, blocks, st, A, jobs, nj);
}

// A hemisphere pass's hand-out tile: texels a wave takes from the launch's
// counter at once.  Default: about 256 rays' worth (1 at 256 samples and
// more, at most 32); one texel per atomic made every wave of a 16-sample
// pass take a texel per 16 rays from the one counter, and the pass ran at a
// third of the 256-sample pass's rays/s.  Thickness takes at least 2: its
// batched hand-out loads a batch's mapping results at once (thickness-256
// on heightfield2048 185 -> 170 ms, on knot4096 4.42 -> 4.16 s).  A tile
// above 1 runs the TB kernels.
uint32_t default_tile(uint32_t S, int kind)
{
    const uint32_t t = S >= 256 ? 1u : std::min(32u, 256u / std::max(S, 1u));
    return kind == FORNOS_SOLVER_THICKNESS ? std::max(t, 2u) : t;
}

uint32_t tile_of(const fornos_solver* s)
{
    return s->tile ? s->tile : default_tile(s->p.sample_count, s->kind);
}

// The autotuners replace a current setting only by one at least 1 % faster.
constexpr float kTuneMargin = 0.99f;

bool launch_shape_ok(uint32_t blocks_per_cu, uint32_t refill, uint32_t stack)
{
    return blocks_per_cu >= 1 && blocks_per_cu <= 16 && refill >= 1 && refill <= 64 &&
           (stack == 16 || stack == 24 || stack == 32);
}

// Device buffers a hemisphere pass writes: the solver's own, or the caller's
// outputs with scratch from the library's fixed array (the no-alloc entry).
struct PassBufs {
    float* out;              // per-texel results (global texel index)
    uint32_t* counts;        // per-texel ray counts (global texel index)
    void* raybuf;            // one launch's BN hit bits or thickness distances
    unsigned int* counter;   // k_rays work counter
    uint32_t chunk_texels;   // texels per k_rays launch (raybuf holds one launch)
    bool timed;              // record per-launch events into s->events
};

PassBufs own_bufs(const fornos_solver* s)
{
    return {s->out, s->counts, s->raybuf, s->counter, s->chunk_texels, s->timed};
}

// Launch a hemisphere pass over texels [t0, t1), or, with a device `list` of
// global texel indices, over list[t0 .. t1).
template <int KIND>
fornos_status launch_hemi(fornos_solver* s, uint32_t t0, uint32_t t1, bool count, const PassBufs& B,
                          const uint32_t* list = nullptr)
{
    fornos_meshmap* mm = s->mm;
    fornos_scene* sc = mm->sc;
    const uint32_t S = s->p.sample_count;
    RayArgs A;
    A.sc = sc->d;
    A.coords = mm->coords;
    A.tidx = mm->tidx;
    A.samples = s->ray_samples;
    const uint32_t mwords = (S + 31) / 32;
    A.cnt = B.counts;
    A.mask = KIND == kBN ? (uint32_t*)B.raybuf : nullptr;
    A.tval = KIND == kThick ? (float*)B.raybuf : nullptr;
    A.counter = B.counter;
    A.stats = s->stats;
    A.fb = count ? s->fbflags : nullptr;
    A.list = list;
    A.S = S;
    A.P = s->P;
    A.mwords = mwords;
    A.mind = s->p.min_distance;
    A.maxd = s->p.max_distance;
    A.refill = s->refill;
    A.tb = tile_of(s);
    ReduceArgs R;
    R.sc = sc->d;
    R.coords = mm->coords;
    R.tidx = mm->tidx;
    R.samples = s->samples;
    R.pixt = mm->pixt;
    R.mask = A.mask;
    R.tval = A.tval;
    R.out = B.out;
    R.counts = B.counts;
    R.list = list;
    R.S = S;
    R.P = s->P;
    R.mwords = mwords;
    R.maxd = s->p.max_distance;
    R.tangent = s->p.tangent_space;
    auto fill = [&](void* p, size_t n, uint32_t v) {
        hipLaunchKernelGGL(k_fill, dim3(div_up(n, kBlock)), dim3(kBlock), 0, sc->stream, (uint32_t*)p, n, v);
    };
    if (count) fill(s->stats, 2 * kStats, 0);
    for (uint32_t c0 = t0; c0 < t1; c0 += B.chunk_texels) {
        const uint32_t nt = std::min(B.chunk_texels, t1 - c0);
        A.t0 = R.t0 = c0;
        A.ntex = R.ntex = nt;
        const uint32_t blocks = std::min<uint32_t>(sc->cus * s->blocks_per_cu, div_up((uint64_t)nt * S, kBlock));
        fill(B.counter, 1, 0);
        if (KIND == kAO && list)
            hipLaunchKernelGGL(k_fill_list, dim3(div_up(nt, kBlock)), dim3(kBlock), 0, sc->stream, B.counts, list + c0, nt, 0u);
        else if (KIND == kAO)
            fill(B.counts + c0, nt, 0);
        if (KIND == kBN) fill(B.raybuf, (size_t)nt * mwords, 0);
        if (KIND == kThick) fill(B.raybuf, (size_t)nt * S, 0x7f7fffffu);  // FLT_MAX
        hipEvent_t* ev = nullptr;
        if (B.timed) {
            s->events.resize(s->events.size() + 3, nullptr);
            ev = &s->events[s->events.size() - 3];
            for (int i = 0; i < 3; ++i) HIP_TRY(hipEventCreate(&ev[i]));
            HIP_TRY(hipEventRecord(ev[0], sc->stream));
        }
        if (count)
            launch_rays<KIND, true, false>(s->stack, blocks, sc->stream, A, nullptr, 0);
        else
            launch_rays<KIND, false, false>(s->stack, blocks, sc->stream, A, nullptr, 0, A.tb > 1);
        if (ev) HIP_TRY(hipEventRecord(ev[1], sc->stream));
        hipLaunchKernelGGL(k_reduce<KIND>, dim3(div_up(nt, kBlock)), dim3(kBlock), 0, sc->stream, R);
        if (ev) HIP_TRY(hipEventRecord(ev[2], sc->stream));
        HIP_TRY(hipGetLastError());
    }
    return FORNOS_OK;
}

// AO solver `s` and its BN partner over texels [t0, t1) (or list[t0 .. t1))
// with one traversal (k_rays<kAOBN>), then both reductions.
fornos_status launch_pair(fornos_solver* s, uint32_t t0, uint32_t t1, const uint32_t* list)
{
    fornos_solver* b = s->partner;
    fornos_meshmap* mm = s->mm;
    fornos_scene* sc = mm->sc;
    const uint32_t S = s->p.sample_count, mwords = (S + 31) / 32;
    RayArgs A;
    A.sc = sc->d;
    A.coords = mm->coords;
    A.tidx = mm->tidx;
    A.samples = s->ray_samples;
    A.cnt = s->counts;
    A.mask = (uint32_t*)b->raybuf;
    A.tval = nullptr;
    A.counter = s->counter;
    A.stats = s->stats;
    A.fb = nullptr;
    A.list = list;
    A.S = S;
    A.P = s->P;
    A.mwords = mwords;
    A.mind = s->p.min_distance;
    A.maxd = s->p.max_distance;
    A.refill = s->refill;
    A.tb = tile_of(s);
    ReduceArgs Ra;
    Ra.sc = sc->d;
    Ra.coords = mm->coords;
    Ra.tidx = mm->tidx;
    Ra.samples = s->samples;
    Ra.pixt = mm->pixt;
    Ra.mask = nullptr;
    Ra.tval = nullptr;
    Ra.out = s->out;
    Ra.counts = s->counts;
    Ra.S = S;
    Ra.P = s->P;
    Ra.mwords = mwords;
    Ra.maxd = s->p.max_distance;
    Ra.tangent = 0;
    Ra.list = list;
    ReduceArgs Rb = Ra;
    Rb.samples = b->samples;
    Rb.mask = A.mask;
    Rb.out = b->out;
    Rb.counts = b->counts;
    Rb.tangent = b->p.tangent_space;
    auto fill = [&](void* p, size_t n, uint32_t v) {
        hipLaunchKernelGGL(k_fill, dim3(div_up(n, kBlock)), dim3(kBlock), 0, sc->stream, (uint32_t*)p, n, v);
    };
    for (uint32_t c0 = t0; c0 < t1; c0 += b->chunk_texels) {
        const uint32_t nt = std::min(b->chunk_texels, t1 - c0);
        A.t0 = Ra.t0 = Rb.t0 = c0;
        A.ntex = Ra.ntex = Rb.ntex = nt;
        const uint32_t blocks = std::min<uint32_t>(sc->cus * s->blocks_per_cu, div_up((uint64_t)nt * S, kBlock));
        fill(s->counter, 1, 0);
        if (list)
            hipLaunchKernelGGL(k_fill_list, dim3(div_up(nt, kBlock)), dim3(kBlock), 0, sc->stream, s->counts, list + c0, nt, 0u);
        else
            fill(s->counts + c0, nt, 0);
        fill(b->raybuf, (size_t)nt * mwords, 0);
        hipEvent_t* ev = nullptr;
        if (s->timed) {
            s->events.resize(s->events.size() + 3, nullptr);
            ev = &s->events[s->events.size() - 3];
            for (int i = 0; i < 3; ++i) HIP_TRY(hipEventCreate(&ev[i]));
            HIP_TRY(hipEventRecord(ev[0], sc->stream));
        }
        launch_rays<kAOBN, false, false>(s->stack, blocks, sc->stream, A, nullptr, 0, A.tb > 1);
        if (ev) HIP_TRY(hipEventRecord(ev[1], sc->stream));
        hipLaunchKernelGGL(k_reduce<kAO>, dim3(div_up(nt, kBlock)), dim3(kBlock), 0, sc->stream, Ra);
        hipLaunchKernelGGL(k_reduce<kBN>, dim3(div_up(nt, kBlock)), dim3(kBlock), 0, sc->stream, Rb);
        if (ev) HIP_TRY(hipEventRecord(ev[2], sc->stream));
        HIP_TRY(hipGetLastError());
    }
    return FORNOS_OK;
}

// A paired AO solver computes its partner's results too while the two are
// at the same point of their passes.
bool fused(const fornos_solver* s) { return s->partner && s->partner->offset == s->offset; }

// Marks the end of the work just queued for `s` (and for its partner when
// the pass computed both), which exports of their results wait for.
fornos_status mark_done(fornos_solver* s)
{
    if (!s->done_ev) HIP_TRY(hipEventCreateWithFlags(&s->done_ev, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(s->done_ev, s->mm->sc->stream));
    return FORNOS_OK;
}

// Work done by a pass of `s` also advances its partner.
fornos_status follow(fornos_solver* s, bool was_fused)
{
    if (fornos_status e = mark_done(s)) return e;
    if (!was_fused) return FORNOS_OK;
    s->partner->offset = s->offset;
    return mark_done(s->partner);
}

// The texels of `mm` sorted by the Morton code of their hit positions, built
// once per mapping result.  Hemisphere passes hand out their rays in this
// order, so the rays in flight leave from a compact patch of the surface
// rather than from a strip along a map row; results do not depend on it.
fornos_status meshmap_order(fornos_meshmap* mm, const uint32_t** out)
{
    *out = nullptr;
    const uint32_t n = mm->n;
    if (n == 0) return FORNOS_OK;
    if (!mm->order_valid) {
        hipStream_t st = mm->sc->stream;
        if (!mm->order)
            if (fornos_status e = dev_alloc(&mm->order, n)) return e;
        uint32_t *keys = nullptr, *keys2 = nullptr, *vals = nullptr;
        void* tmp = nullptr;
        size_t tmp_bytes = 0;
        fornos_status e = dev_alloc(&keys, n);
        if (!e) e = dev_alloc(&keys2, n);
        if (!e) e = dev_alloc(&vals, n);
        if (!e && hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys2, vals, mm->order, (int)n, 0, 32, st) !=
                      hipSuccess)
            e = set_error(FORNOS_HIP_ERROR, "texel order: radix sort size query failed");
        if (!e) e = dev_alloc((uint8_t**)&tmp, std::max<size_t>(tmp_bytes, 1));
        if (!e) {
            hipLaunchKernelGGL(k_order_keys, dim3(div_up(n, kBlock)), dim3(kBlock), 0, st, mm->sc->d, mm->coords, mm->tidx,
                               n, keys, vals);
            if (hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys2, vals, mm->order, (int)n, 0, 32, st) !=
                hipSuccess)
                e = set_error(FORNOS_HIP_ERROR, "texel order: radix sort failed");
        }
        if (hipStreamSynchronize(st) != hipSuccess && !e) e = set_error(FORNOS_HIP_ERROR, "texel order: stream failed");
        dev_free(keys); dev_free(keys2); dev_free(vals); dev_free((uint8_t*)tmp);
        if (e) return e;
        mm->order_valid = true;
    }
    *out = mm->order;
    return FORNOS_OK;
}

// Diagonal of the scene's box (the reference tree's root).
float scene_diagonal(const fornos_scene* sc)
{
    if (sc->bvh.n_nodes == 0) return 0.0f;
    float lo[3], hi[3];
    std::memcpy(lo, &sc->bvh.nodes[0], sizeof(lo));
    std::memcpy(hi, &sc->bvh.nodes[4], sizeof(hi));
    double d = 0;
    for (int k = 0; k < 3; ++k) d += (double)(hi[k] - lo[k]) * (hi[k] - lo[k]);
    return (float)std::sqrt(d);
}

// Rays at least this fraction of the scene's diagonal long cross the scene:
// the texels of one map row then send rays through most of it, and handing
// them out by the Morton code of their origins keeps what the rays in flight
// touch within the caches (knot4096, maxD 10 over a 5-unit scene: AO -10 %,
// thickness -20 %).  Shorter rays stay near their origins, and a map row's
// texels already lie close together (heightfield2048, maxD 0.2 over 1.4: the
// paired pass is 3 % slower in Morton order).
constexpr float kMortonRayReach = 0.25f;

// FORNOS_RAY_ORDER_TEXELS or _MORTON: the order a solver's passes use.
int effective_order(const fornos_solver* s)
{
    if (!s->hemi() || s->ray_order == FORNOS_RAY_ORDER_TEXELS) return FORNOS_RAY_ORDER_TEXELS;
    if (s->ray_order == FORNOS_RAY_ORDER_AUTO && !(s->p.max_distance >= kMortonRayReach * scene_diagonal(s->mm->sc)))
        return FORNOS_RAY_ORDER_TEXELS;
    return FORNOS_RAY_ORDER_MORTON;
}

// The order a solver's passes hand out their texels in (null: the texel
// set's own order).
fornos_status pass_order(fornos_solver* s, const uint32_t** list)
{
    *list = nullptr;
    if (effective_order(s) == FORNOS_RAY_ORDER_TEXELS) return FORNOS_OK;
    return meshmap_order(s->mm, list);
}

fornos_status solver_launch(fornos_solver* s, uint32_t t0, uint32_t t1, bool count, const uint32_t* list = nullptr)
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
    case FORNOS_SOLVER_AO: return launch_hemi<kAO>(s, t0, t1, count, own_bufs(s), list);
    case FORNOS_SOLVER_BENT_NORMALS: return launch_hemi<kBN>(s, t0, t1, count, own_bufs(s), list);
    case FORNOS_SOLVER_THICKNESS: return launch_hemi<kThick>(s, t0, t1, count, own_bufs(s), list);
    }
    HIP_TRY(hipGetLastError());
    return FORNOS_OK;
}

// A hemisphere solver's two device copies of its sample table, as float4
// rows: t4 in table order (w = 0) and r4 in ray order, each permutation's
// samples sorted by the Morton code of their disk position (x, y) so that
// consecutive rays of a texel -- the lanes of a wave -- point in similar
// directions, with w carrying the sample's index, under which its result is
// stored.  Both depend only on the table, so the last few small ones are
// kept on the host (compared word for word, not by hash): the solvers of an
// AO + bent-normal pair and of the next bake on the same table skip the
// conversion and the sort.
struct RayTable {
    uint32_t S = 0, P = 0;
    uint64_t hash = 0;         // of the table's words (solvers compare tables by it)
    std::vector<float> table;  // 3 * P * S, as given
    std::vector<float4> t4, r4;
};

std::shared_ptr<const RayTable> ray_table(const float* table, uint32_t S, uint32_t P)
{
    // at most 2 kept, of at most 2^20 samples each (44 bytes a sample)
    constexpr size_t kKeep = 2, kMaxSamples = (size_t)1 << 20;
    static std::mutex mu;
    static std::vector<std::shared_ptr<const RayTable>> kept;  // most recent last
    const size_t n = (size_t)P * S;
    if (n <= kMaxSamples) {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t i = 0; i < kept.size(); ++i) {
            const std::shared_ptr<const RayTable> k = kept[i];
            if (k->S == S && k->P == P && std::memcmp(k->table.data(), table, sizeof(float) * 3 * n) == 0) {
                kept.erase(kept.begin() + i);
                kept.push_back(k);
                return k;
            }
        }
    }
    auto rt = std::make_shared<RayTable>();
    rt->S = S;
    rt->P = P;
    rt->t4.resize(n);
    rt->r4.resize(n);
    {
        // FNV-1a over the table's 32-bit words in blocks of 2^16 words
        // (hashed on several threads), then over the block hashes in order
        const size_t words = 3 * n, block = 1u << 16;
        std::vector<uint64_t> bh((words + block - 1) / block);
        parallel_chunks(bh.size(), [&](size_t b0, size_t b1) {
            for (size_t b = b0; b < b1; ++b) {
                uint64_t h = 1469598103934665603ull;
                for (size_t i = b * block; i < std::min(words, (b + 1) * block); ++i) {
                    uint32_t w;
                    std::memcpy(&w, &table[i], 4);
                    h = (h ^ w) * 1099511628211ull;
                }
                bh[b] = h;
            }
        }, 1);
        uint64_t h = 1469598103934665603ull;
        for (uint64_t x : bh) h = (h ^ x) * 1099511628211ull;
        rt->hash = h;
    }
    auto spread = [](uint32_t v) {  // 16 bits -> even bit positions
        v &= 0xffffu;
        v = (v | (v << 8)) & 0x00ff00ffu;
        v = (v | (v << 4)) & 0x0f0f0f0fu;
        v = (v | (v << 2)) & 0x33333333u;
        v = (v | (v << 1)) & 0x55555555u;
        return v;
    };
    auto quant = [](float x) {
        const float u = std::min(1.0f, std::max(0.0f, 0.5f * (x + 1.0f)));
        return (uint32_t)(u * 65535.0f);
    };
    // (key << 32 | index): unique, so an unstable sort orders them as a
    // stable sort of the keys would; threads only where the work pays for
    // starting them
    parallel_chunks(P, [&](size_t p0, size_t p1) {
        std::vector<uint64_t> key(S);
        for (size_t p = p0; p < p1; ++p) {
            float4* t4 = &rt->t4[p * S];
            const float* row = table + 3 * p * S;
            for (uint32_t i = 0; i < S; ++i) {
                t4[i] = make_float4(row[3 * i], row[3 * i + 1], row[3 * i + 2], 0.0f);
                key[i] = (uint64_t)(spread(quant(t4[i].x)) | (spread(quant(t4[i].y)) << 1)) << 32 | i;
            }
            std::sort(key.begin(), key.end());
            for (uint32_t j = 0; j < S; ++j) {
                const uint32_t i = (uint32_t)key[j];
                const float4 v = t4[i];
                rt->r4[p * S + j] = make_float4(v.x, v.y, v.z, __builtin_bit_cast(float, i));
            }
        }
    }, std::max<size_t>(1, ((size_t)1 << 17) / std::max<uint32_t>(S, 1)));
    if (n <= kMaxSamples) {
        rt->table.assign(table, table + 3 * n);
        std::lock_guard<std::mutex> lk(mu);
        kept.push_back(rt);
        if (kept.size() > kKeep) kept.erase(kept.begin());
    }
    return rt;
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
    sc->cus = (uint32_t)std::max(1, cus);
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
    CullBVH cull, cull_ao;
    {
        std::vector<float> sp4((size_t)4 * spos.size());
        std::memcpy(sp4.data(), spos.data(), sizeof(float) * sp4.size());
        build_cull_bvh(B, sp4, amax * (1.0f / 65536.0f), cull);
        // AO's own tree with leaves split into triangle groups, built when the
        // surface-area estimate predicts a gain and kept when the built tree
        // confirms it (elongated leaves, as on heightfields, gain most;
        // compact ones gain nothing and only deepen the tree)
        if (cull.sah_split_pred < kSplitPredict * cull.sah) {
            build_cull_bvh(B, sp4, amax * (1.0f / 65536.0f), cull_ao, kSplitCost, false);
            if (!(cull_ao.sah < kSplitKeep * cull.sah)) cull_ao = CullBVH();
        }
        // the other passes keep whole leaves; on such scenes group boxes
        // inside them still spare triangle tests
        if (cull_ao.n_wide) add_leaf_groups(cull, sp4, amax * (1.0f / 65536.0f));
    }
    const std::vector<uint32_t> chain = reference_chain_links(B);
    const double build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count();

    HIP_TRY(hipStreamCreateWithFlags(&sc->stream, hipStreamNonBlocking));
    if (fornos_status st = dev_alloc(&sc->cnodes, (size_t)2 * cull.n_nodes)) return st;
    if (fornos_status st = dev_alloc(&sc->cref, cull.ref_leaf.size())) return st;
    if (fornos_status st = dev_alloc(&sc->chain, chain.size())) return st;
    HIP_TRY(hipMemcpy(sc->cnodes, cull.nodes.data(), sizeof(uint32_t) * 8 * (size_t)cull.n_nodes, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(sc->cref, cull.ref_leaf.data(), sizeof(uint32_t) * cull.ref_leaf.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(sc->chain, chain.data(), sizeof(uint32_t) * chain.size(), hipMemcpyHostToDevice));
    sc->d.cnodes = sc->cnodes;
    sc->d.cref = sc->cref;
    sc->d.chain = sc->chain;
    sc->d.n_cnodes = cull.n_nodes;
    if (fornos_status st = dev_alloc(&sc->wide, (size_t)8 * cull.n_wide)) return st;
    if (fornos_status st = dev_alloc(&sc->wleaf, cull.wide_leaves.size() / 4)) return st;
    HIP_TRY(hipMemcpy(sc->wide, cull.wide.data(), sizeof(uint32_t) * cull.wide.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(sc->wleaf, cull.wide_leaves.data(), sizeof(uint32_t) * cull.wide_leaves.size(), hipMemcpyHostToDevice));
    sc->d.wide = sc->wide;
    sc->d.wleaf = sc->wleaf;
    sc->d.wide_ao = sc->wide;
    sc->d.wleaf_ao = sc->wleaf;
    if (fornos_status st = dev_alloc(&sc->wsub, (size_t)8 * std::max<uint32_t>(1, cull.n_sub))) return st;
    if (cull.n_sub)
        HIP_TRY(hipMemcpy(sc->wsub, cull.sub.data(), sizeof(uint32_t) * cull.sub.size(), hipMemcpyHostToDevice));
    sc->d.wsub = sc->wsub;
    if (cull_ao.n_wide) {
        if (fornos_status st = dev_alloc(&sc->wide_ao, (size_t)8 * cull_ao.n_wide)) return st;
        if (fornos_status st = dev_alloc(&sc->wleaf_ao, cull_ao.wide_leaves.size() / 4)) return st;
        HIP_TRY(hipMemcpy(sc->wide_ao, cull_ao.wide.data(), sizeof(uint32_t) * cull_ao.wide.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(sc->wleaf_ao, cull_ao.wide_leaves.data(), sizeof(uint32_t) * cull_ao.wide_leaves.size(),
                          hipMemcpyHostToDevice));
        sc->d.wide_ao = sc->wide_ao;
        sc->d.wleaf_ao = sc->wleaf_ao;
    }
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
                            sizeof(uint32_t) * (chain.size() + cull.ref_leaf.size() + cull.wide.size() +
                                                cull.wide_leaves.size() + cull_ao.wide.size() +
                                                cull_ao.wide_leaves.size() + cull.sub.size());
    sc->info.ao_groups = cull_ao.n_groups;
    sc->info.leaf_group_nodes = cull.n_sub;
    sc->info.sah = cull.sah;
    sc->info.sah_ao = cull_ao.n_wide ? cull_ao.sah : cull.sah;
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

void fornos_scene_destroy(fornos_scene* sc) { release(sc); }

fornos_status fornos_meshmap_create(fornos_scene* sc, const fornos_texel_view* tv, int cull,
                                    fornos_meshmap** out)
{
    if (!sc || !tv || !out) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_create: null argument");
    *out = nullptr;
    if (tv->count > 0 && (!tv->positions || !tv->directions))
        return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_create: texel positions/directions missing");
    std::unique_ptr<fornos_meshmap> mm(new fornos_meshmap());
    mm->sc = sc;
    retain(sc);
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
    if (tv->indices) {
        if (fornos_status st = dev_alloc(&mm->indices, mm->n)) return st;
        HIP_TRY(hipMemcpy(mm->indices, tv->indices, sizeof(uint32_t) * mm->n, hipMemcpyHostToDevice));
    }
    if (fornos_status st = dev_alloc(&mm->coords, mm->n)) return st;
    if (fornos_status st = dev_alloc(&mm->tidx, mm->n)) return st;
    *out = mm.release();
    return FORNOS_OK;
}

fornos_status fornos_meshmap_create_from_mesh(fornos_scene* sc, const fornos_mesh_view* lo,
                                              const fornos_mesh_view* dirs, uint32_t W, uint32_t H, int cull,
                                              fornos_meshmap** out)
{
    return fornos_meshmap_create_from_mesh_hybrid(sc, lo, dirs, W, H, 0.0f, cull, out);
}

fornos_status fornos_meshmap_create_from_mesh_hybrid(fornos_scene* sc, const fornos_mesh_view* lo,
                                                     const fornos_mesh_view* dirs, uint32_t W, uint32_t H,
                                                     float edge, int cull, fornos_meshmap** out)
{
    if (!sc || !lo || !out) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_create_from_mesh: null argument");
    *out = nullptr;
    if (W == 0 || H == 0 || W > 16384 || H > 16384)
        return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_create_from_mesh: map size must be in [1, 16384]");
    if (dirs && dirs->n_vertices != lo->n_vertices)
        return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_create_from_mesh: mapping mesh must be a copy of the low-poly mesh");
    if (!(edge >= 0.0f)) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_create_from_mesh_hybrid: edge must be >= 0");
    const bool hybrid = edge > 0.0f;
    // corner records of rasterTriangle (Src/compute.cpp:85-126)
    const uint32_t nt = lo->n_triangles;
    std::vector<float> tris((size_t)kTriFloats * nt, 0.0f);
    for (uint32_t t = 0; t < nt; ++t) {
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = lo->triangles[3 * t + k];
            if (v >= lo->n_vertices) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_create_from_mesh: bad vertex index");
            const uint32_t pi = lo->vertices[3 * v], ti = lo->vertices[3 * v + 1], ni = lo->vertices[3 * v + 2];
            if (ti == UINT32_MAX || ni == UINT32_MAX || !lo->texcoords || !lo->normals)
                return set_error(FORNOS_INVALID_ARG, "Low poly mesh is missing texture coordinates or normals information");
            if (pi >= lo->n_positions || ti >= lo->n_texcoords || ni >= lo->n_normals)
                return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_create_from_mesh: vertex index out of range");
            float* c = &tris[(size_t)t * kTriFloats + k * kCornerFloats];
            std::memcpy(c, &lo->positions[3 * (size_t)pi], 12);
            std::memcpy(c + 3, &lo->texcoords[2 * (size_t)ti], 8);
            std::memcpy(c + 5, &lo->normals[3 * (size_t)ni], 12);
            const float* d = &lo->normals[3 * (size_t)ni];
            if (dirs) {
                const uint32_t di = dirs->vertices[3 * v + 2];
                if (di >= dirs->n_normals) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_create_from_mesh: bad mapping normal index");
                d = &dirs->normals[3 * (size_t)di];
            }
            std::memcpy(c + 8, d, 12);
            if (lo->tangents && lo->bitangents) {
                // the Hybrid raster reads the tangent frame by normal index (Src/compute.cpp:264-269)
                const uint32_t tv = hybrid ? ni : v;
                if (tv >= lo->n_vertices)
                    return set_error(FORNOS_INVALID_ARG, "Hybrid mapping: a normal index is out of the tangent frame's range");
                std::memcpy(c + 11, &lo->tangents[3 * (size_t)tv], 12);
                std::memcpy(c + 14, &lo->bitangents[3 * (size_t)tv], 12);
            }
        }
        if (hybrid) {
            V2 u[3];
            for (int k = 0; k < 3; ++k) u[k] = {tris[(size_t)t * kTriFloats + k * kCornerFloats + 3],
                                                tris[(size_t)t * kTriFloats + k * kCornerFloats + 4]};
            if (!texel_box(u, W, H, true).ok)
                return set_error(FORNOS_INVALID_ARG, "texture coordinates too far outside [0, 1] for the Hybrid mapping");
        }
    }
    std::unique_ptr<fornos_meshmap> mm(new fornos_meshmap());
    mm->sc = sc;
    retain(sc);
    mm->width = W;
    mm->height = H;
    mm->cull = cull ? 1 : 0;
    hipStream_t st = sc->stream;
    const uint64_t npx = (uint64_t)W * H;
    struct Tmp {
        float* tris = nullptr;
        float* map = nullptr;
        unsigned long long* owner = nullptr;
        uint32_t *flags = nullptr, *slot = nullptr;
        void* scan = nullptr;
        ~Tmp() { dev_free(tris); dev_free(map); dev_free(owner); dev_free(flags); dev_free(slot); dev_free(scan); }
    } tmp;
    if (fornos_status s = dev_alloc(&tmp.tris, tris.size())) return s;
    if (fornos_status s = dev_alloc(&tmp.map, 15 * npx)) return s;
    if (fornos_status s = dev_alloc(&tmp.owner, npx)) return s;
    if (fornos_status s = dev_alloc(&tmp.flags, npx)) return s;
    if (fornos_status s = dev_alloc(&tmp.slot, npx)) return s;
    HIP_TRY(hipMemcpyAsync(tmp.tris, tris.data(), sizeof(float) * tris.size(), hipMemcpyHostToDevice, st));
    const uint32_t pblocks = div_up(npx, kBlock);
    hipLaunchKernelGGL(k_fill, dim3(div_up(2 * npx, kBlock)), dim3(kBlock), 0, st, (uint32_t*)tmp.owner, (size_t)(2 * npx), 0u);
    hipLaunchKernelGGL(k_fill, dim3(div_up(15 * npx, kBlock)), dim3(kBlock), 0, st, (uint32_t*)tmp.map, (size_t)(15 * npx), 0u);
    RasterArgs R{tmp.tris, nt, W, H, edge, tmp.owner, tmp.map};
    if (nt) hipLaunchKernelGGL(k_raster_bid, dim3(nt), dim3(kBlock), 0, st, R);
    hipLaunchKernelGGL(k_raster_write, dim3(pblocks), dim3(kBlock), 0, st, R);
    hipLaunchKernelGGL(k_texel_flags, dim3(pblocks), dim3(kBlock), 0, st, tmp.map + 3 * npx, npx, tmp.flags);
    size_t scan_bytes = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, tmp.flags, tmp.slot, (int)npx, st));
    uint8_t* sb = nullptr;
    if (fornos_status s = dev_alloc(&sb, scan_bytes)) return s;
    tmp.scan = sb;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.scan, scan_bytes, tmp.flags, tmp.slot, (int)npx, st));
    uint32_t last_slot = 0, last_flag = 0;
    HIP_TRY(hipMemcpyAsync(&last_slot, tmp.slot + npx - 1, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&last_flag, tmp.flags + npx - 1, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    mm->n = last_slot + last_flag;
    mm->work_count = div_up(mm->n, kGroup) * kGroup;
    if (fornos_status s = dev_alloc(&mm->pix, (size_t)2 * mm->n)) return s;
    if (fornos_status s = dev_alloc(&mm->pixt, (size_t)3 * mm->n)) return s;
    if (fornos_status s = dev_alloc(&mm->indices, mm->n)) return s;
    if (fornos_status s = dev_alloc(&mm->coords, mm->n)) return s;
    if (fornos_status s = dev_alloc(&mm->tidx, mm->n)) return s;
    hipLaunchKernelGGL(k_texel_compact, dim3(pblocks), dim3(kBlock), 0, st, tmp.map, npx, tmp.flags, tmp.slot,
                       mm->pix, mm->pixt, mm->indices);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    *out = mm.release();
    return FORNOS_OK;
}

fornos_status fornos_meshmap_read_texels(const fornos_meshmap* mm, float* p, float* d, float* n, float* t, float* b,
                                         uint32_t* indices)
{
    if (!mm) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_read_texels: null handle");
    HIP_TRY(hipStreamSynchronize(mm->sc->stream));
    if (mm->n == 0) return FORNOS_OK;
    std::vector<float4> buf((size_t)3 * mm->n);
    auto unpack = [&](float* dst, size_t stride, size_t off) {
        if (!dst) return;
        for (size_t i = 0; i < mm->n; ++i) {
            const float4 v = buf[stride * i + off];
            dst[3 * i] = v.x; dst[3 * i + 1] = v.y; dst[3 * i + 2] = v.z;
        }
    };
    if (p || d) {
        HIP_TRY(hipMemcpy(buf.data(), mm->pix, sizeof(float4) * 2 * mm->n, hipMemcpyDeviceToHost));
        unpack(p, 2, 0);
        unpack(d, 2, 1);
    }
    if (n || t || b) {
        if (!mm->pixt) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_read_texels: the texel set has no tangent frames");
        HIP_TRY(hipMemcpy(buf.data(), mm->pixt, sizeof(float4) * 3 * mm->n, hipMemcpyDeviceToHost));
        unpack(n, 3, 0);
        unpack(t, 3, 1);
        unpack(b, 3, 2);
    }
    if (indices) {
        if (!mm->indices) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_read_texels: the texel set has no MapUV indices");
        HIP_TRY(hipMemcpy(indices, mm->indices, sizeof(uint32_t) * mm->n, hipMemcpyDeviceToHost));
    }
    return FORNOS_OK;
}

static fornos_status meshmap_launch(fornos_meshmap* mm, uint32_t t0, uint32_t t1)
{
    if (t1 <= t0) return FORNOS_OK;
    mm->order_valid = false;
    const uint32_t blocks = div_up(t1 - t0, kBlock);
    hipStream_t st = mm->sc->stream;
    const DevScene& d = mm->sc->d;
    const bool wt = mm->predicate == FORNOS_PREDICATE_WATERTIGHT;
    if (!mm->redo) {
        if (fornos_status e = dev_alloc(&mm->redo, 2)) return e;
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, st, mm->redo, (size_t)2, 0u);
    }
    if (!wt && mm->traversal == FORNOS_TRAVERSAL_WIDE && d.n_cnodes > 0) {
        if (mm->cull)
            hipLaunchKernelGGL((k_meshmap_wide<true>), dim3(blocks), dim3(kBlock), 0, st, mm->pix, t0, t1, d, mm->coords,
                               mm->tidx, mm->redo, mm->wide_stack);
        else
            hipLaunchKernelGGL((k_meshmap_wide<false>), dim3(blocks), dim3(kBlock), 0, st, mm->pix, t0, t1, d, mm->coords,
                               mm->tidx, mm->redo, mm->wide_stack);
    } else if (mm->cull && wt) {
        hipLaunchKernelGGL((k_meshmap<true, true>), dim3(blocks), dim3(kBlock), 0, st, mm->pix, t0, t1, d, mm->coords, mm->tidx);
    } else if (mm->cull) {
        hipLaunchKernelGGL((k_meshmap<true, false>), dim3(blocks), dim3(kBlock), 0, st, mm->pix, t0, t1, d, mm->coords, mm->tidx);
    } else if (wt) {
        hipLaunchKernelGGL((k_meshmap<false, true>), dim3(blocks), dim3(kBlock), 0, st, mm->pix, t0, t1, d, mm->coords, mm->tidx);
    } else {
        hipLaunchKernelGGL((k_meshmap<false, false>), dim3(blocks), dim3(kBlock), 0, st, mm->pix, t0, t1, d, mm->coords, mm->tidx);
    }
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

fornos_status fornos_meshmap_set_predicate(fornos_meshmap* mm, int predicate)
{
    if (!mm || (predicate != FORNOS_PREDICATE_REFERENCE && predicate != FORNOS_PREDICATE_WATERTIGHT))
        return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_set_predicate: FORNOS_PREDICATE_REFERENCE or _WATERTIGHT");
    mm->predicate = predicate;
    return FORNOS_OK;
}

fornos_status fornos_meshmap_reset(fornos_meshmap* mm)
{
    if (!mm) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_reset: null handle");
    mm->offset = 0;
    mm->marks.clear();
    if (mm->redo) {
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, mm->sc->stream, mm->redo, (size_t)2, 0u);
        HIP_TRY(hipGetLastError());
    }
    return FORNOS_OK;
}

fornos_status fornos_meshmap_set_traversal(fornos_meshmap* mm, int traversal, uint32_t stack)
{
    if (!mm || (traversal != FORNOS_TRAVERSAL_WIDE && traversal != FORNOS_TRAVERSAL_REFERENCE))
        return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_set_traversal: FORNOS_TRAVERSAL_WIDE or _REFERENCE");
    if (stack != 0 && (stack < 3 || stack > (uint32_t)kMapStack))
        return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_set_traversal: stack 0 (default) or 3..24 entries");
    mm->traversal = traversal;
    mm->wide_stack = stack ? stack : (uint32_t)kMapStack;
    return FORNOS_OK;
}

fornos_status fornos_meshmap_redo_counts(fornos_meshmap* mm, uint32_t* undecided, uint32_t* overflow)
{
    if (!mm || !undecided || !overflow) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_redo_counts: null argument");
    uint32_t h[2] = {0, 0};
    HIP_TRY(hipStreamSynchronize(mm->sc->stream));
    if (mm->redo) HIP_TRY(hipMemcpy(h, mm->redo, sizeof(h), hipMemcpyDeviceToHost));
    *undecided = h[0];
    *overflow = h[1];
    return FORNOS_OK;
}

fornos_status fornos_meshmap_progress(const fornos_meshmap* mm, float* out)
{
    if (!mm || !out) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_progress: null argument");
    *out = mm->work_count ? (float)mm->offset / (float)mm->work_count : 1.0f;
    return FORNOS_OK;
}

// Submission without waiting (the runner's pacing): the remaining texels in
// `parts` launches, an event after each; poll() reports finished texels.
static fornos_status poll_marks(Marks& marks, hipStream_t st, uint32_t from, uint32_t n, bool submitted,
                                float* progress, int* done)
{
    if (marks.empty()) {
        const hipError_t r = hipStreamQuery(st);
        if (r != hipErrorNotReady) HIP_TRY(r);
        const bool idle = r == hipSuccess;
        if (progress) *progress = submitted ? 1.0f : 0.0f;
        if (done) *done = submitted && idle;
        return FORNOS_OK;
    }
    uint64_t fin = 0;
    if (fornos_status e = marks.poll(&fin)) return e;
    fin = std::max<uint64_t>(fin, from);
    if (progress) *progress = marks.complete() ? 1.0f : (n ? (float)fin / (float)n : 1.0f);
    if (done) *done = marks.complete();
    return FORNOS_OK;
}

fornos_status fornos_meshmap_submit(fornos_meshmap* mm, uint32_t parts)
{
    if (!mm) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_submit: null handle");
    if (parts == 0 || parts > 4096) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_submit: parts must be in [1, 4096]");
    mm->marks.clear();
    const uint32_t t0 = std::min(mm->offset, mm->n);
    mm->mark_from = t0;
    const uint32_t per = std::max<uint32_t>(1, div_up(mm->n - t0, parts));
    for (uint32_t a = t0; a < mm->n; a += per) {
        const uint32_t b = std::min(mm->n, a + per);
        if (fornos_status st = meshmap_launch(mm, a, b)) return st;
        if (fornos_status st = mm->marks.add(mm->sc->stream, b)) return st;
    }
    mm->offset = mm->work_count;
    return FORNOS_OK;
}

fornos_status fornos_meshmap_poll(fornos_meshmap* mm, float* progress, int* done)
{
    if (!mm) return set_error(FORNOS_INVALID_ARG, "fornos_meshmap_poll: null handle");
    return poll_marks(mm->marks, mm->sc->stream, mm->mark_from, mm->n, mm->offset >= mm->work_count, progress, done);
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

void fornos_meshmap_destroy(fornos_meshmap* mm) { release(mm); }

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
    retain(mm);
    if ((kind == FORNOS_SOLVER_NORMALS || kind == FORNOS_SOLVER_BENT_NORMALS) && params->tangent_space && !mm->pixt)
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_create: tangent space needs per-texel normals, tangents and bitangents");
    if (s->hemi()) {
        const uint32_t S = params->sample_count;
        if (S == 0 || S > 4096) return set_error(FORNOS_INVALID_ARG, "fornos_solver_create: sample_count must be in [1, 4096]");
        if (!table || n_perm == 0) return set_error(FORNOS_INVALID_ARG, "fornos_solver_create: sample table missing");
        s->P = n_perm;
        s->work_total = (uint64_t)mm->work_count * S;
        const std::shared_ptr<const RayTable> rt = ray_table(table, S, n_perm);
        s->table_hash = rt->hash;
        if (fornos_status st = dev_alloc(&s->samples, rt->t4.size())) return st;
        HIP_TRY(hipMemcpy(s->samples, rt->t4.data(), sizeof(float4) * rt->t4.size(), hipMemcpyHostToDevice));
        if (fornos_status st = dev_alloc(&s->ray_samples, rt->r4.size())) return st;
        HIP_TRY(hipMemcpy(s->ray_samples, rt->r4.data(), sizeof(float4) * rt->r4.size(), hipMemcpyHostToDevice));
        if (fornos_status st = dev_alloc(&s->counts, mm->n)) return st;
        if (fornos_status st = dev_alloc(&s->stats, kStats)) return st;
        if (fornos_status st = dev_alloc(&s->counter, 1)) return st;
        // per-ray results of one k_rays launch of at most 2^28 rays: hit bits
        // (BN, 32 MiB) or t values (thickness, 1 GiB); AO counts in place
        s->chunk_texels = std::max<uint32_t>(1, (1u << 28) / S);
        const size_t tex = std::min<uint32_t>(s->chunk_texels, std::max<uint32_t>(1, mm->n));
        if (kind != FORNOS_SOLVER_AO) {
            const size_t words = kind == FORNOS_SOLVER_THICKNESS ? tex * S : tex * ((S + 31) / 32);
            uint32_t* rb = nullptr;
            if (fornos_status st = dev_alloc(&rb, words)) return st;
            s->raybuf = rb;
        }
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
            const bool fz = fused(s);
            const uint32_t* list = nullptr;
            if (fornos_status st = pass_order(s, &list)) return st;
            if (fornos_status st = fz ? launch_pair(s, std::min(t0, n), t1, list)
                                      : solver_launch(s, std::min(t0, n), t1, false, list))
                return st;
            s->offset += work;
            if (fornos_status st = follow(s, fz)) return st;
        } else {
            const uint64_t work = std::min<uint64_t>(s->work_total - s->offset, kWorkPerStep);
            const uint32_t t0 = (uint32_t)s->offset, t1 = (uint32_t)std::min<uint64_t>(s->offset + work, n);
            if (fornos_status st = solver_launch(s, std::min(t0, n), t1, false)) return st;
            s->offset += work;
            if (fornos_status st = mark_done(s)) return st;
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
    const uint32_t* list = nullptr;
    if (fornos_status st = pass_order(s, &list)) return st;
    s->timed = s->hemi();
    const bool fz = fused(s);
    const fornos_status st = fz ? launch_pair(s, t0, n, list) : solver_launch(s, t0, n, false, list);
    s->timed = false;
    if (st) return st;
    s->offset = s->work_total;
    if (fornos_status st2 = follow(s, fz)) return st2;
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

fornos_status fornos_solver_run_texels(fornos_solver* s, const uint32_t* texels, uint32_t n, float* ms)
{
    if (!s || (!texels && n)) return set_error(FORNOS_INVALID_ARG, "fornos_solver_run_texels: null argument");
    if (!s->hemi())
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_run_texels: not a hemisphere solver (AO, bent normals, thickness)");
    const uint32_t count = s->mm->n;
    {
        // distinct and in range: a texel listed twice would count its rays twice
        std::vector<uint64_t> seen(((size_t)count + 63) / 64, 0);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t t = texels[i];
            if (t >= count || (seen[t >> 6] >> (t & 63)) & 1)
                return set_error(FORNOS_INVALID_ARG, "fornos_solver_run_texels: texel indices must be distinct and below "
                                                     "the texel count");
            seen[t >> 6] |= 1ull << (t & 63);
        }
    }
    if (s->owner) return set_error(FORNOS_INVALID_ARG, "fornos_solver_run_texels: a paired bent-normals solver runs through its AO solver");
    hipStream_t st = s->mm->sc->stream;
    if (n > s->dlist_cap) {
        HIP_TRY(hipStreamSynchronize(st));
        dev_free(s->dlist);
        s->dlist = nullptr;
        s->dlist_cap = 0;
        if (fornos_status e = dev_alloc(&s->dlist, n)) return e;
        s->dlist_cap = n;
    }
    Timer tm(st);
    if (fornos_status e = tm.start()) return e;
    if (n) {
        HIP_TRY(hipMemcpyAsync(s->dlist, texels, sizeof(uint32_t) * n, hipMemcpyHostToDevice, st));
        fornos_status e = FORNOS_OK;
        if (s->partner) {
            e = launch_pair(s, 0, n, s->dlist);
        } else {
            switch (s->kind) {
            case FORNOS_SOLVER_AO: e = launch_hemi<kAO>(s, 0, n, false, own_bufs(s), s->dlist); break;
            case FORNOS_SOLVER_BENT_NORMALS: e = launch_hemi<kBN>(s, 0, n, false, own_bufs(s), s->dlist); break;
            default: e = launch_hemi<kThick>(s, 0, n, false, own_bufs(s), s->dlist); break;
            }
        }
        if (e) return e;
    }
    if (fornos_status e = mark_done(s)) return e;
    if (s->partner)
        if (fornos_status e = mark_done(s->partner)) return e;
    return tm.stop(ms);
}

fornos_status fornos_solver_run_all_noalloc(fornos_solver* s, float* out, uint32_t* counts)
{
    if (!s || !out || !counts) return set_error(FORNOS_INVALID_ARG, "fornos_solver_run_all_noalloc: null argument");
    if (!s->hemi())
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_run_all_noalloc: not a hemisphere solver (AO, bent normals, thickness)");
    const uint32_t n = s->mm->n, S = s->p.sample_count;
    // one texel's share of a launch's per-ray results: thickness distances or bent-normal hit bits
    const size_t per_texel = s->kind == FORNOS_SOLVER_THICKNESS ? 4ull * S
                             : s->kind == FORNOS_SOLVER_BENT_NORMALS ? 4ull * ((S + 31) / 32) : 0;
    if (per_texel > kScratchData)
        return set_error(FORNOS_UNSUPPORTED, "fornos_solver_run_all_noalloc: one texel's per-ray results exceed the "
                                             "64 MiB scratch array");
    if (n == 0) return FORNOS_OK;
    uint8_t* base = nullptr;
    if (fornos_status st = scratch_for(s->mm->sc->device, &base)) return st;
    PassBufs B{out, counts, base + kScratchHead, (unsigned int*)base, s->chunk_texels, false};
    if (per_texel) B.chunk_texels = (uint32_t)std::min<size_t>(B.chunk_texels, kScratchData / per_texel);
    fornos_status st = FORNOS_OK;
    switch (s->kind) {
    case FORNOS_SOLVER_AO: st = launch_hemi<kAO>(s, 0, n, false, B); break;
    case FORNOS_SOLVER_BENT_NORMALS: st = launch_hemi<kBN>(s, 0, n, false, B); break;
    default: st = launch_hemi<kThick>(s, 0, n, false, B); break;
    }
    if (st) return st;
    HIP_TRY(hipStreamSynchronize(s->mm->sc->stream));
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

fornos_status fornos_solver_submit(fornos_solver* s, uint32_t parts)
{
    if (!s) return set_error(FORNOS_INVALID_ARG, "fornos_solver_submit: null handle");
    if (parts == 0 || parts > 4096) return set_error(FORNOS_INVALID_ARG, "fornos_solver_submit: parts must be in [1, 4096]");
    s->marks.clear();
    const uint32_t n = s->mm->n;
    const uint32_t S = s->hemi() ? s->p.sample_count : 1;
    const uint32_t t0 = (uint32_t)std::min<uint64_t>(s->offset / S, n);
    s->mark_from = t0;
    const bool fz = fused(s);
    const uint32_t* list = nullptr;
    if (fornos_status st = pass_order(s, &list)) return st;
    const uint32_t per = std::max<uint32_t>(1, div_up(n - t0, parts));
    for (uint32_t a = t0; a < n; a += per) {
        const uint32_t b = std::min(n, a + per);
        if (fornos_status st = fz ? launch_pair(s, a, b, list) : solver_launch(s, a, b, false, list)) return st;
        if (fornos_status st = s->marks.add(s->mm->sc->stream, b)) return st;
    }
    s->offset = s->work_total;
    return follow(s, fz);
}

fornos_status fornos_solver_poll(fornos_solver* s, float* progress, int* done)
{
    if (!s) return set_error(FORNOS_INVALID_ARG, "fornos_solver_poll: null handle");
    return poll_marks(s->marks, s->mm->sc->stream, s->mark_from, s->mm->n, s->offset >= s->work_total, progress,
                      done);
}

fornos_status fornos_solver_pair(fornos_solver* ao, fornos_solver* bn)
{
    if (!ao) return set_error(FORNOS_INVALID_ARG, "fornos_solver_pair: null handle");
    if (ao->kind != FORNOS_SOLVER_AO) return set_error(FORNOS_INVALID_ARG, "fornos_solver_pair: first solver must be AO");
    if (!bn) {
        if (ao->partner) ao->partner->owner = nullptr;
        ao->partner = nullptr;
        return FORNOS_OK;
    }
    if (bn->kind != FORNOS_SOLVER_BENT_NORMALS)
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_pair: second solver must be bent normals");
    if (bn->mm != ao->mm) return set_error(FORNOS_INVALID_ARG, "fornos_solver_pair: solvers on different mesh mappings");
    if (bn->p.sample_count != ao->p.sample_count || bn->P != ao->P || bn->table_hash != ao->table_hash)
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_pair: different sample tables");
    if (__builtin_bit_cast(uint32_t, bn->p.min_distance) != __builtin_bit_cast(uint32_t, ao->p.min_distance) ||
        __builtin_bit_cast(uint32_t, bn->p.max_distance) != __builtin_bit_cast(uint32_t, ao->p.max_distance))
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_pair: different min / max distances");
    if ((bn->owner && bn->owner != ao) || (ao->partner && ao->partner != bn))
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_pair: a solver is already paired");
    ao->partner = bn;
    bn->owner = ao;
    return FORNOS_OK;
}

fornos_status fornos_solver_reset(fornos_solver* s)
{
    if (!s) return set_error(FORNOS_INVALID_ARG, "fornos_solver_reset: null handle");
    s->offset = 0;
    s->marks.clear();
    if (s->partner) {  // a pair is reset together, so it runs fused again
        s->partner->offset = 0;
        s->partner->marks.clear();
    }
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

fornos_status fornos_solver_count_work(fornos_solver* s, uint64_t* nodes, uint64_t* tris, uint64_t* rays,
                                       uint64_t* overflows)
{
    if (!s) return set_error(FORNOS_INVALID_ARG, "fornos_solver_count_work: null handle");
    if (!s->hemi()) return set_error(FORNOS_INVALID_ARG, "fornos_solver_count_work: not a hemisphere solver");
    if (!s->fbflags && s->mm->n) {
        if (fornos_status st = dev_alloc(&s->fbflags, s->mm->n)) return st;
    }
    if (s->mm->n)
        hipLaunchKernelGGL(k_fill, dim3(div_up(s->mm->n, kBlock)), dim3(kBlock), 0, s->mm->sc->stream, s->fbflags,
                           (size_t)s->mm->n, 0u);
    const uint32_t* list = nullptr;
    if (fornos_status st = pass_order(s, &list)) return st;
    if (fornos_status st = solver_launch(s, 0, s->mm->n, true, list)) return st;
    unsigned long long h[kStats] = {};
    HIP_TRY(hipStreamSynchronize(s->mm->sc->stream));
    HIP_TRY(hipMemcpy(h, s->stats, sizeof(h), hipMemcpyDeviceToHost));
    if (nodes) *nodes = h[0];
    if (tris) *tris = h[1];
    if (rays) *rays = h[2];
    if (overflows) *overflows = h[3];
    return FORNOS_OK;
}

fornos_status fornos_solver_work_counters(const fornos_solver* s, uint64_t* counters, uint32_t n)
{
    if (!s || !counters || n == 0 || n > kStats)
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_work_counters: 1 to " + std::to_string(kStats) + " counters");
    if (!s->hemi()) return set_error(FORNOS_INVALID_ARG, "fornos_solver_work_counters: not a hemisphere solver");
    HIP_TRY(hipStreamSynchronize(s->mm->sc->stream));
    HIP_TRY(hipMemcpy(counters, s->stats, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    return FORNOS_OK;
}

fornos_status fornos_solver_fallback_flags(const fornos_solver* s, uint32_t* flags)
{
    if (!s || !flags) return set_error(FORNOS_INVALID_ARG, "fornos_solver_fallback_flags: null argument");
    if (!s->hemi()) return set_error(FORNOS_INVALID_ARG, "fornos_solver_fallback_flags: not a hemisphere solver");
    if (!s->mm->n) return FORNOS_OK;
    if (!s->fbflags) return set_error(FORNOS_INVALID_ARG, "fornos_solver_fallback_flags: no counting run yet (fornos_solver_count_work)");
    HIP_TRY(hipStreamSynchronize(s->mm->sc->stream));
    HIP_TRY(hipMemcpy(flags, s->fbflags, sizeof(uint32_t) * s->mm->n, hipMemcpyDeviceToHost));
    return FORNOS_OK;
}

fornos_status fornos_solver_set_launch(fornos_solver* s, uint32_t blocks_per_cu, uint32_t refill, uint32_t stack)
{
    if (!s || !launch_shape_ok(blocks_per_cu, refill, stack))
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_set_launch: blocks_per_cu in [1,16], refill in [1,64], stack 16, 24 or 32");
    s->blocks_per_cu = blocks_per_cu;
    s->refill = refill;
    s->stack = stack;
    return FORNOS_OK;
}

fornos_status fornos_solver_set_tile(fornos_solver* s, uint32_t texels)
{
    if (!s || texels > 32) return set_error(FORNOS_INVALID_ARG, "fornos_solver_set_tile: texels in [0, 32] (0: default)");
    if (!s->hemi())
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_set_tile: not a hemisphere solver (AO, bent normals, thickness)");
    s->tile = texels;
    return FORNOS_OK;
}

fornos_status fornos_solver_get_tile(const fornos_solver* s, uint32_t* texels)
{
    if (!s || !texels) return set_error(FORNOS_INVALID_ARG, "fornos_solver_get_tile: null argument");
    if (!s->hemi())
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_get_tile: not a hemisphere solver (AO, bent normals, thickness)");
    *texels = tile_of(s);
    return FORNOS_OK;
}

fornos_status fornos_solver_autotune(fornos_solver* s, float fraction, uint32_t* best_bpc, uint32_t* best_refill,
                                     uint32_t* best_stack, float* best_ms)
{
    if (!s || !(fraction > 0.0f) || fraction > 1.0f)
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_autotune: fraction must be in (0, 1]");
    if (!s->hemi())
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_autotune: not a hemisphere solver (AO, bent normals, thickness)");
    if (s->owner)
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_autotune: a paired bent-normals solver is tuned through its AO solver");
    const uint32_t n = s->mm->n, S = s->p.sample_count;
    // the first `fraction` of the texels in the pass's own order, at least
    // 2^22 rays, so the launch reaches its steady state
    const uint32_t m = std::min<uint32_t>(n, std::max<uint32_t>(div_up(1u << 22, S), (uint32_t)((double)n * fraction)));
    const uint32_t* list = nullptr;
    if (fornos_status e = pass_order(s, &list)) return e;
    hipStream_t st = s->mm->sc->stream;
    const uint32_t keep[4] = {s->blocks_per_cu, s->refill, s->stack, s->tile};
    const int keep_order = s->ray_order;
    uint32_t cur[4] = {keep[0], keep[1], keep[2], tile_of(s)};  // blocks per CU, refill, stack, tile
    // time a launch shape over `count` texels of `lst` (null: texels [0, count))
    auto time_shape = [&](const uint32_t* shape, const uint32_t* lst, uint32_t count, float* ms) -> fornos_status {
        s->blocks_per_cu = shape[0];
        s->refill = shape[1];
        s->stack = shape[2];
        s->tile = shape[3];
        for (int rep = 0; rep < 2; ++rep) {  // the first run warms the caches
            Timer tm(st);
            if (fornos_status e = tm.start()) return e;
            const fornos_status e = s->partner ? launch_pair(s, 0, count, lst) : solver_launch(s, 0, count, false, lst);
            if (e) return e;
            if (fornos_status e2 = tm.stop(ms)) return e2;
        }
        return FORNOS_OK;
    };
    // coordinate descent: stack depth, then blocks per CU, then refill
    // threshold, then the hand-out tile (around its default: 1 or 2 texels
    // at 256 samples and more; thickness 1, 2 or 4)
    const uint32_t t0 = default_tile(S, s->kind);
    std::vector<uint32_t> tiles = {std::max(1u, t0 / 2), t0, std::min(32u, 2 * t0)};
    tiles.erase(std::unique(tiles.begin(), tiles.end()), tiles.end());
    const std::vector<uint32_t> cand[4] = {{4, 6, 8, 12}, {16, 24, 32, 40, 48}, {16, 24, 32}, tiles};
    // On each axis the current value is timed first and another replaces it
    // only when at least kTuneMargin faster: the prefix's timings vary by
    // about 1 %, and near-equal shapes chosen on noise ran the whole pass
    // slower (heightfield2048 AO-256: refill 40 for 32, 185.4 -> 187.6 ms).
    float best = FLT_MAX;
    fornos_status err = FORNOS_OK;
    if (m > 0) {
        for (int axis : {2, 0, 1, 3}) {
            uint32_t pick = cur[axis];
            float inc = 0;
            if ((err = time_shape(cur, list, m, &inc))) break;
            float axis_best = inc;
            for (uint32_t v : cand[axis]) {
                if (v == cur[axis]) continue;
                uint32_t shape[4] = {cur[0], cur[1], cur[2], cur[3]};
                shape[axis] = v;
                float ms = 0;
                if ((err = time_shape(shape, list, m, &ms))) break;
                if (ms < kTuneMargin * inc && ms < axis_best) {
                    axis_best = ms;
                    pick = v;
                }
            }
            if (err) break;
            cur[axis] = pick;
            best = axis_best;
        }
    }
    // the ray order (the one sort of the rays: none, or by the Morton code of
    // their texels' hit positions once per mapping result), timed on the same
    // texels in both orders -- every k-th texel of the set, so that each
    // order's rays in flight spread as they do in the whole pass
    int order = effective_order(s);
    if (!err && m > 0) {
        const uint32_t* mort = nullptr;
        err = meshmap_order(s->mm, &mort);
        std::vector<uint32_t> perm(n), lists;
        if (!err && hipMemcpy(perm.data(), mort, sizeof(uint32_t) * n, hipMemcpyDeviceToHost) != hipSuccess)
            err = set_error(FORNOS_HIP_ERROR, "fornos_solver_autotune: reading the texel order failed");
        const uint32_t k = std::max<uint32_t>(1, n / m);
        if (!err) {
            for (uint32_t g = 0; g < n; g += k) lists.push_back(g);
            const size_t cnt = lists.size();
            for (uint32_t g : perm)
                if (g % k == 0) lists.push_back(g);
            uint32_t* dl = nullptr;
            if (lists.size() != 2 * cnt) err = set_error(FORNOS_HIP_ERROR, "fornos_solver_autotune: texel order is not a permutation");
            if (!err) err = dev_alloc(&dl, lists.size());
            if (!err && hipMemcpy(dl, lists.data(), sizeof(uint32_t) * lists.size(), hipMemcpyHostToDevice) != hipSuccess)
                err = set_error(FORNOS_HIP_ERROR, "fornos_solver_autotune: uploading the sample texels failed");
            float ms_t = 0, ms_m = 0;
            if (!err) err = time_shape(cur, dl, (uint32_t)cnt, &ms_t);
            if (!err) err = time_shape(cur, dl + cnt, (uint32_t)cnt, &ms_m);
            if (hipStreamSynchronize(st) != hipSuccess && !err) err = set_error(FORNOS_HIP_ERROR, "fornos_solver_autotune: stream failed");
            dev_free(dl);
            // the automatic order is the incumbent, replaced only when the
            // other is at least kTuneMargin faster
            if (!err) {
                const bool morton = order == FORNOS_RAY_ORDER_MORTON;
                if (morton ? ms_t < kTuneMargin * ms_m : ms_m < kTuneMargin * ms_t)
                    order = morton ? FORNOS_RAY_ORDER_TEXELS : FORNOS_RAY_ORDER_MORTON;
            }
        }
    }
    s->blocks_per_cu = err ? keep[0] : cur[0];
    s->refill = err ? keep[1] : cur[1];
    s->stack = err ? keep[2] : cur[2];
    s->tile = err ? keep[3] : cur[3];
    s->ray_order = err || m == 0 ? keep_order : order;
    if (err) return err;
    if (best_bpc) *best_bpc = cur[0];
    if (best_refill) *best_refill = cur[1];
    if (best_stack) *best_stack = cur[2];
    if (best_ms) *best_ms = best;
    return FORNOS_OK;
}

fornos_status fornos_solver_set_ray_order(fornos_solver* s, int order)
{
    if (!s || (order != FORNOS_RAY_ORDER_TEXELS && order != FORNOS_RAY_ORDER_MORTON && order != FORNOS_RAY_ORDER_AUTO))
        return set_error(FORNOS_INVALID_ARG, "fornos_solver_set_ray_order: FORNOS_RAY_ORDER_TEXELS, _MORTON or _AUTO");
    s->ray_order = order;
    return FORNOS_OK;
}

fornos_status fornos_solver_get_ray_order(const fornos_solver* s, int* order)
{
    if (!s || !order) return set_error(FORNOS_INVALID_ARG, "fornos_solver_get_ray_order: null argument");
    *order = s->ray_order;
    return FORNOS_OK;
}

fornos_status fornos_solver_effective_ray_order(const fornos_solver* s, int* order)
{
    if (!s || !order) return set_error(FORNOS_INVALID_ARG, "fornos_solver_effective_ray_order: null argument");
    *order = effective_order(s);
    return FORNOS_OK;
}

fornos_status fornos_solver_get_launch(const fornos_solver* s, uint32_t* blocks_per_cu, uint32_t* refill,
                                       uint32_t* stack)
{
    if (!s) return set_error(FORNOS_INVALID_ARG, "fornos_solver_get_launch: null handle");
    if (blocks_per_cu) *blocks_per_cu = s->blocks_per_cu;
    if (refill) *refill = s->refill;
    if (stack) *stack = s->stack;
    return FORNOS_OK;
}

void fornos_solver_destroy(fornos_solver* s) { release(s); }

}  // extern "C"

struct fornos_batch {
    std::vector<fornos_solver*> solvers;  // one reference held on each
    std::vector<RayArgs> jobs;
    hipStream_t stream = nullptr;
    RayArgs* djobs = nullptr;
    uint32_t* counters = nullptr;
    int cus = 256;
    uint32_t blocks_per_cu = 8, refill = kRefill, stack = kStack;
    std::vector<uint32_t> base_tile;  // per job: its solver's hand-out tile
    int tile_shift = 0;               // the jobs' tiles are base_tile halved (-1), kept (0) or doubled (+1)
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    ~fornos_batch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        dev_free(djobs);
        dev_free(counters);
        if (stream) (void)hipStreamDestroy(stream);
        for (fornos_solver* s : solvers) release(s);
    }
};

namespace {

// Enqueues one batched pass: counters and AO counts cleared, one k_rays over
// the job table (jobs[i].ntex texels each), one k_reduce per job.
fornos_status batch_enqueue(fornos_batch* b, const std::vector<RayArgs>& jobs, RayArgs* djobs)
{
    hipStream_t st = b->stream;
    const uint32_t nj = (uint32_t)jobs.size();
    hipLaunchKernelGGL(k_fill, dim3(div_up(nj, kBlock)), dim3(kBlock), 0, st, b->counters, (size_t)nj, 0u);
    for (const RayArgs& J : jobs) {
        if (!J.ntex) continue;
        if (J.list)
            hipLaunchKernelGGL(k_fill_list, dim3(div_up(J.ntex, kBlock)), dim3(kBlock), 0, st, J.cnt, J.list, J.ntex, 0u);
        else
            hipLaunchKernelGGL(k_fill, dim3(div_up(J.ntex, kBlock)), dim3(kBlock), 0, st, J.cnt, (size_t)J.ntex, 0u);
    }
    uint64_t rays = 0;
    bool tb = false;  // some job's hand-out tile is above 1: the TB kernel
    for (const RayArgs& J : jobs) {
        rays += (uint64_t)J.ntex * J.S;
        tb |= J.tb > 1;
    }
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((uint64_t)b->cus * b->blocks_per_cu, div_up(rays, kBlock));
    if (blocks) launch_rays<kAO, false, true>(b->stack, blocks, st, jobs[0], djobs, nj, tb);
    for (uint32_t j = 0; j < nj; ++j) {
        const fornos_solver* s = b->solvers[j];
        ReduceArgs R{};
        R.sc = jobs[j].sc;
        R.tidx = jobs[j].tidx;
        R.out = s->out;
        R.counts = s->counts;
        R.S = jobs[j].S;
        R.t0 = 0;
        R.ntex = jobs[j].ntex;
        R.list = jobs[j].list;
        if (R.ntex) hipLaunchKernelGGL(k_reduce<kAO>, dim3(div_up(R.ntex, kBlock)), dim3(kBlock), 0, st, R);
    }
    HIP_TRY(hipGetLastError());
    return FORNOS_OK;
}

void batch_drop_graph(fornos_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// A job's hand-out tile under the batch's tile shift.
uint32_t shifted_tile(uint32_t base, int shift)
{
    return shift < 0 ? std::max(1u, base >> 1) : shift > 0 ? std::min(32u, base << 1) : base;
}

}  // namespace

extern "C" {

fornos_status fornos_batch_create(fornos_solver* const* solvers, uint32_t n, fornos_batch** out)
{
    if (!solvers || !out || n == 0) return set_error(FORNOS_INVALID_ARG, "fornos_batch_create: no solvers");
    *out = nullptr;
    std::unique_ptr<fornos_batch> b(new fornos_batch());
    for (uint32_t j = 0; j < n; ++j) {
        fornos_solver* s = solvers[j];
        if (!s || s->kind != FORNOS_SOLVER_AO)
            return set_error(FORNOS_UNSUPPORTED, "fornos_batch_create: every job must be an AO solver");
        if ((uint64_t)s->mm->n * s->p.sample_count >= (1ull << 31))
            return set_error(FORNOS_UNSUPPORTED, "fornos_batch_create: a job has 2^31 rays or more");
        b->solvers.push_back(s);
        retain(s);
    }
    // longest first (texels x samples x log2 triangles): the persistent waves
    // start on the big jobs and the small ones fill the last waves' tail
    auto cost = [](const fornos_solver* s) {
        return (double)s->mm->n * s->p.sample_count * std::log2(2.0 + s->mm->sc->info.n_triangles);
    };
    std::stable_sort(b->solvers.begin(), b->solvers.end(),
                     [&](const fornos_solver* x, const fornos_solver* y) { return cost(x) > cost(y); });
    HIP_TRY(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->solvers[0]->mm->sc->device));
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    if (fornos_status st = dev_alloc(&b->counters, n)) return st;
    if (fornos_status st = dev_alloc(&b->djobs, n)) return st;
    for (uint32_t j = 0; j < n; ++j) {
        fornos_solver* s = b->solvers[j];
        RayArgs A{};
        A.sc = s->mm->sc->d;
        A.coords = s->mm->coords;
        A.tidx = s->mm->tidx;
        A.samples = s->ray_samples;
        A.cnt = s->counts;
        A.counter = b->counters + j;
        A.stats = s->stats;
        A.S = s->p.sample_count;
        A.P = s->P;
        A.t0 = 0;
        A.ntex = s->mm->n;
        A.mwords = (A.S + 31) / 32;
        A.mind = s->p.min_distance;
        A.maxd = s->p.max_distance;
        A.refill = b->refill;
        const uint32_t* list = nullptr;
        if (fornos_status st = pass_order(s, &list)) return st;
        A.list = list;
        A.tb = tile_of(s);
        b->base_tile.push_back(A.tb);
        b->jobs.push_back(A);
    }
    HIP_TRY(hipMemcpy(b->djobs, b->jobs.data(), sizeof(RayArgs) * n, hipMemcpyHostToDevice));
    *out = b.release();
    return FORNOS_OK;
}

fornos_status fornos_batch_set_launch(fornos_batch* b, uint32_t blocks_per_cu, uint32_t refill, uint32_t stack)
{
    if (!b || !launch_shape_ok(blocks_per_cu, refill, stack))
        return set_error(FORNOS_INVALID_ARG, "fornos_batch_set_launch: blocks_per_cu in [1,16], refill in [1,64], stack 16, 24 or 32");
    if (blocks_per_cu == b->blocks_per_cu && refill == b->refill && stack == b->stack) return FORNOS_OK;
    b->blocks_per_cu = blocks_per_cu;
    b->refill = refill;
    b->stack = stack;
    for (RayArgs& A : b->jobs) A.refill = refill;
    HIP_TRY(hipMemcpy(b->djobs, b->jobs.data(), sizeof(RayArgs) * b->jobs.size(), hipMemcpyHostToDevice));
    batch_drop_graph(b);
    return FORNOS_OK;
}

fornos_status fornos_batch_get_launch(const fornos_batch* b, uint32_t* blocks_per_cu, uint32_t* refill,
                                      uint32_t* stack)
{
    if (!b) return set_error(FORNOS_INVALID_ARG, "fornos_batch_get_launch: null handle");
    if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu;
    if (refill) *refill = b->refill;
    if (stack) *stack = b->stack;
    return FORNOS_OK;
}

fornos_status fornos_batch_run(fornos_batch* b, float* ms)
{
    if (!b) return set_error(FORNOS_INVALID_ARG, "fornos_batch_run: null handle");
    if (!b->exec) {
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const fornos_status st = batch_enqueue(b, b->jobs, b->djobs);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (st) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        HIP_TRY(e);
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    Timer tm(b->stream);
    if (fornos_status st = tm.start()) return st;
    HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    if (fornos_status st = tm.stop(ms)) return st;
    for (fornos_solver* s : b->solvers) s->offset = s->work_total;
    return FORNOS_OK;
}

fornos_status fornos_batch_autotune(fornos_batch* b, float fraction, uint32_t* best_bpc, uint32_t* best_refill,
                                    uint32_t* best_stack, float* best_ms)
{
    if (!b || !(fraction > 0.0f) || fraction > 1.0f)
        return set_error(FORNOS_INVALID_ARG, "fornos_batch_autotune: fraction must be in (0, 1]");
    // a prefix of each job's texels (row-major: a band of the map): the
    // fraction, but at least 2^22 rays, so each job reaches its steady state
    std::vector<RayArgs> sample = b->jobs;
    for (RayArgs& A : sample) {
        const uint32_t floor_tex = std::min<uint32_t>(A.ntex, div_up(1u << 22, A.S));
        A.ntex = std::max<uint32_t>(floor_tex, (uint32_t)(A.ntex * fraction));
    }
    RayArgs* dsample = nullptr;
    if (fornos_status st = dev_alloc(&dsample, sample.size())) return st;
    struct Free { RayArgs* p; ~Free() { dev_free(p); } } guard{dsample};
    const uint32_t keep[3] = {b->blocks_per_cu, b->refill, b->stack};
    // blocks per CU, refill, stack, tile shift + 1
    uint32_t cur[4] = {keep[0], keep[1], keep[2], (uint32_t)(b->tile_shift + 1)};
    auto time_shape = [&](const uint32_t* shape, float* ms) -> fornos_status {
        b->blocks_per_cu = shape[0];
        b->stack = shape[2];
        for (size_t j = 0; j < sample.size(); ++j) {
            sample[j].refill = shape[1];
            sample[j].tb = shifted_tile(b->base_tile[j], (int)shape[3] - 1);
        }
        HIP_TRY(hipMemcpy(dsample, sample.data(), sizeof(RayArgs) * sample.size(), hipMemcpyHostToDevice));
        for (int rep = 0; rep < 2; ++rep) {  // the first run warms the caches
            Timer tm(b->stream);
            if (fornos_status st = tm.start()) return st;
            if (fornos_status st = batch_enqueue(b, sample, dsample)) return st;
            if (fornos_status st = tm.stop(ms)) return st;
        }
        return FORNOS_OK;
    };
    // coordinate descent: stack depth, then blocks per CU, then refill
    // threshold, then the jobs' hand-out tiles (halved, kept, doubled)
    // (the current value of each axis kept unless another is kTuneMargin faster)
    const std::vector<uint32_t> cand[4] = {{4, 6, 8, 12}, {16, 24, 32, 40, 48}, {16, 24, 32}, {0, 1, 2}};
    float best = FLT_MAX;
    for (int axis : {2, 0, 1, 3}) {
        uint32_t pick = cur[axis];
        float inc = 0;
        if (fornos_status st = time_shape(cur, &inc)) return st;
        float axis_best = inc;
        for (uint32_t v : cand[axis]) {
            if (v == cur[axis]) continue;
            uint32_t shape[4] = {cur[0], cur[1], cur[2], cur[3]};
            shape[axis] = v;
            float ms = 0;
            if (fornos_status st = time_shape(shape, &ms)) return st;
            if (ms < kTuneMargin * inc && ms < axis_best) {
                axis_best = ms;
                pick = v;
            }
        }
        cur[axis] = pick;
        best = axis_best;
    }
    b->blocks_per_cu = keep[0];
    b->refill = keep[1];
    b->stack = keep[2];
    b->tile_shift = (int)cur[3] - 1;
    for (size_t j = 0; j < b->jobs.size(); ++j) b->jobs[j].tb = shifted_tile(b->base_tile[j], b->tile_shift);
    HIP_TRY(hipMemcpy(b->djobs, b->jobs.data(), sizeof(RayArgs) * b->jobs.size(), hipMemcpyHostToDevice));
    if (fornos_status st = fornos_batch_set_launch(b, cur[0], cur[1], cur[2])) return st;
    batch_drop_graph(b);
    if (best_bpc) *best_bpc = cur[0];
    if (best_refill) *best_refill = cur[1];
    if (best_stack) *best_stack = cur[2];
    if (best_ms) *best_ms = best;
    return FORNOS_OK;
}

fornos_status fornos_batch_set_tile_shift(fornos_batch* b, int shift)
{
    if (!b || shift < -1 || shift > 1) return set_error(FORNOS_INVALID_ARG, "fornos_batch_set_tile_shift: shift -1, 0 or 1");
    if (shift == b->tile_shift) return FORNOS_OK;
    b->tile_shift = shift;
    for (size_t j = 0; j < b->jobs.size(); ++j) b->jobs[j].tb = shifted_tile(b->base_tile[j], shift);
    HIP_TRY(hipMemcpy(b->djobs, b->jobs.data(), sizeof(RayArgs) * b->jobs.size(), hipMemcpyHostToDevice));
    batch_drop_graph(b);
    return FORNOS_OK;
}

fornos_status fornos_batch_get_tile_shift(const fornos_batch* b, int* shift)
{
    if (!b || !shift) return set_error(FORNOS_INVALID_ARG, "fornos_batch_get_tile_shift: null argument");
    *shift = b->tile_shift;
    return FORNOS_OK;
}

fornos_status fornos_batch_get_tile(const fornos_batch* b, const fornos_solver* s, uint32_t* tile)
{
    if (!b || !s || !tile) return set_error(FORNOS_INVALID_ARG, "fornos_batch_get_tile: null argument");
    for (size_t j = 0; j < b->solvers.size(); ++j)
        if (b->solvers[j] == s) {
            *tile = b->jobs[j].tb;
            return FORNOS_OK;
        }
    return set_error(FORNOS_INVALID_ARG, "fornos_batch_get_tile: the solver is not a job of this batch");
}

void fornos_batch_destroy(fornos_batch* b) { delete b; }

}  // extern "C"

// ---------------------------------------------------------------------------
// Denoiser input / output on the texel set (include/fornos_denoise.h)
// ---------------------------------------------------------------------------
namespace {

__device__ inline float finite_or0(float v) { return isfinite(v) ? v : 0.0f; }

struct GatherArgs {
    const uint32_t* indices;
    const float *ao, *bn, *pos, *nrm, *hgt;  // pos, nrm, hgt null: from the mapping (below)
    const float4 *coords, *spos, *snrm;      // the mapping's results and the scene's triangle soups
    const uint32_t* tidx;
    float* image;
    uint32_t n, npix;
    float cx, cy, cz, inv_extent, dist_scale;
};

// MAPPED: position, object-space normal and mapping distance computed here
// from the mapping, as the position, normals and height passes compute them
// (texel_attr, texel_height), rather than read from those solvers' results
template <bool MAPPED>
__global__ void k_dn_gather(GatherArgs A)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= A.n) return;
    const uint32_t pix = A.indices[q];
    if (pix >= A.npix) return;
    float* o = A.image + (size_t)pix * FORNOS_DN_IN;
    V3 p, nr;
    float h;
    if (MAPPED) {
        p = texel_attr<false>(A.coords, A.tidx, A.spos, q);
        nr = texel_attr<true>(A.coords, A.tidx, A.snrm, q);
        h = texel_height(A.coords, q);
    } else {
        p = mk(A.pos[3 * (size_t)q], A.pos[3 * (size_t)q + 1], A.pos[3 * (size_t)q + 2]);
        nr = mk(A.nrm[3 * (size_t)q], A.nrm[3 * (size_t)q + 1], A.nrm[3 * (size_t)q + 2]);
        h = A.hgt[q];
    }
    // the pixel's 12 channels as three 16-byte stores (48-byte records)
    static_assert(FORNOS_DN_IN == 12 && FORNOS_DN_MASK_CHANNEL == 11, "the gather writes 12 channels, mask last");
    float4* o4 = reinterpret_cast<float4*>(o);
    o4[0] = make_float4(finite_or0(A.ao[q]), finite_or0(A.bn[3 * (size_t)q]), finite_or0(A.bn[3 * (size_t)q + 1]),
                        finite_or0(A.bn[3 * (size_t)q + 2]));
    o4[1] = make_float4(finite_or0((p.x - A.cx) * A.inv_extent), finite_or0((p.y - A.cy) * A.inv_extent),
                        finite_or0((p.z - A.cz) * A.inv_extent), finite_or0(nr.x));
    o4[2] = make_float4(finite_or0(nr.y), finite_or0(nr.z), finite_or0(h * A.dist_scale), 1.0f);
}

__global__ void k_dn_fill(float* p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// ao[q] = image AO clamped to [0, 1]; bn[q] = normalize(image bent normal)
__global__ void k_dn_scatter(const uint32_t* indices, const float* image, uint32_t n, uint32_t npix, float* ao,
                             float* bn)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint32_t pix = indices[q];
    if (pix >= npix) return;
    const float* v = image + (size_t)pix * FORNOS_DN_OUT;
    ao[q] = fminf(fmaxf(v[0], 0.0f), 1.0f);
    const float x = v[1], y = v[2], z = v[3];
    const float l = sqrtf(x * x + y * y + z * z);
    const float s = l > 0.0f ? 1.0f / l : 0.0f;
    bn[3 * (size_t)q + 0] = x * s;
    bn[3 * (size_t)q + 1] = y * s;
    bn[3 * (size_t)q + 2] = z * s;
}

// Per tile of tile x tile pixels: the mean of (noisy AO - denoised AO) over
// the tile's texels against its standard error under the binomial noise of
// an S-sample estimate, var = p (1 - p) / S per texel.  A tile fails when
// |mean| > z * sqrt(sum var) / count + slack.
__global__ void k_dn_tiles(const float* input, const float* output, uint32_t w, uint32_t h, uint32_t tile,
                           uint32_t samples, float z, float slack, uint32_t* fail)
{
    const uint32_t tx = blockIdx.x, ty = blockIdx.y;
    float sum = 0.0f, var = 0.0f, cnt = 0.0f;
    for (uint32_t i = threadIdx.x; i < tile * tile; i += blockDim.x) {
        const uint32_t x = tx * tile + i % tile, y = ty * tile + i / tile;
        if (x >= w || y >= h) continue;
        const size_t p = (size_t)y * w + x;
        if (!(input[p * FORNOS_DN_IN + FORNOS_DN_MASK_CHANNEL] > 0.5f)) continue;
        const float d = fminf(fmaxf(output[p * FORNOS_DN_OUT], 0.0f), 1.0f);
        sum += input[p * FORNOS_DN_IN] - d;
        var += d * (1.0f - d) / (float)samples;
        cnt += 1.0f;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off);
        var += __shfl_xor(var, off);
        cnt += __shfl_xor(cnt, off);
    }
    __shared__ float red[3][4];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wv] = sum;
        red[1][wv] = var;
        red[2][wv] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float S = 0, V = 0, N = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            S += red[0][k];
            V += red[1][k];
            N += red[2][k];
        }
        const bool bad = N > 0.0f && fabsf(S / N) > z * sqrtf(V) / N + slack;
        fail[(size_t)ty * gridDim.x + tx] = bad ? 1u : 0u;
    }
}

fornos_status dn_solver_ok(const fornos_solver* s, const fornos_meshmap* mm, int kind, const char* what)
{
    if (!s || s->mm != mm || s->kind != kind)
        return set_error(FORNOS_INVALID_ARG, std::string("fornos_denoise_gather: ") + what +
                                                 " must be a solver of that kind on the same mesh mapping");
    return FORNOS_OK;
}

}  // namespace

extern "C" {

fornos_status fornos_denoise_gather(const fornos_meshmap* mm, const fornos_solver* ao, const fornos_solver* bn,
                                    const fornos_solver* position, const fornos_solver* normals,
                                    const fornos_solver* height, const float* center, float inv_extent,
                                    float dist_scale, float* image)
{
    if (!mm || !center || !image) return set_error(FORNOS_INVALID_ARG, "fornos_denoise_gather: null argument");
    if (reinterpret_cast<uintptr_t>(image) % 16)
        return set_error(FORNOS_INVALID_ARG, "fornos_denoise_gather: image must be 16-byte aligned");
    if (!mm->indices) return set_error(FORNOS_INVALID_ARG, "fornos_denoise_gather: the texel set has no map indices");
    if (fornos_status s = dn_solver_ok(ao, mm, FORNOS_SOLVER_AO, "ao")) return s;
    if (fornos_status s = dn_solver_ok(bn, mm, FORNOS_SOLVER_BENT_NORMALS, "bn")) return s;
    const bool mapped = !position && !normals && !height;
    if (mapped) {
        if (mm->offset < mm->work_count)
            return set_error(FORNOS_INVALID_ARG, "fornos_denoise_gather: the mesh mapping has not been run");
    } else {
        if (fornos_status s = dn_solver_ok(position, mm, FORNOS_SOLVER_POSITION, "position")) return s;
        if (fornos_status s = dn_solver_ok(normals, mm, FORNOS_SOLVER_NORMALS, "normals")) return s;
        if (fornos_status s = dn_solver_ok(height, mm, FORNOS_SOLVER_HEIGHT, "height")) return s;
    }
    hipStream_t st = mm->sc->stream;
    const size_t npix = (size_t)mm->width * mm->height;
    hipLaunchKernelGGL(k_dn_fill, dim3(div_up(npix * FORNOS_DN_IN, kBlock)), dim3(kBlock), 0, st, image,
                       npix * FORNOS_DN_IN);
    GatherArgs A{mm->indices, ao->out, bn->out, nullptr, nullptr, nullptr, mm->coords, mm->sc->spos, mm->sc->snrm,
                 mm->tidx, image, mm->n, (uint32_t)npix, center[0], center[1], center[2], inv_extent, dist_scale};
    if (!mapped) {
        A.pos = position->out;
        A.nrm = normals->out;
        A.hgt = height->out;
    }
    if (mm->n) {
        if (mapped)
            hipLaunchKernelGGL(k_dn_gather<true>, dim3(div_up(mm->n, kBlock)), dim3(kBlock), 0, st, A);
        else
            hipLaunchKernelGGL(k_dn_gather<false>, dim3(div_up(mm->n, kBlock)), dim3(kBlock), 0, st, A);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return FORNOS_OK;
}

fornos_status fornos_denoise_scatter(const fornos_meshmap* mm, const float* image, float* ao, float* bn)
{
    if (!mm || !image || !ao || !bn) return set_error(FORNOS_INVALID_ARG, "fornos_denoise_scatter: null argument");
    if (!mm->indices) return set_error(FORNOS_INVALID_ARG, "fornos_denoise_scatter: the texel set has no map indices");
    hipStream_t st = mm->sc->stream;
    if (mm->n)
        hipLaunchKernelGGL(k_dn_scatter, dim3(div_up(mm->n, kBlock)), dim3(kBlock), 0, st, (const uint32_t*)mm->indices,
                           image, mm->n, mm->width * mm->height, ao, bn);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return FORNOS_OK;
}

fornos_status fornos_denoise_check_tiles(const float* input, const float* output, uint32_t width, uint32_t height,
                                         uint32_t tile, uint32_t samples, float z, float slack, uint32_t* fail)
{
    if (!input || !output || !fail || tile == 0 || samples == 0 || width == 0 || height == 0)
        return set_error(FORNOS_INVALID_ARG, "fornos_denoise_check_tiles: bad argument");
    const dim3 grid(div_up(width, tile), div_up(height, tile));
    hipLaunchKernelGGL(k_dn_tiles, grid, dim3(kBlock), 0, nullptr, input, output, width, height, tile, samples, z,
                       slack, fail);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return FORNOS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Image export on the device (include/fornos_export.h)
// ---------------------------------------------------------------------------
namespace {

// One export on its scene's export stream, holding the scene's export lock:
// waits for the work queued on the scene stream so far, carves its device
// buffers from the scene's workspace (grown when too small) and returns its
// results through stream-ordered copies.
struct Export {
    fornos_scene* sc;
    std::unique_lock<std::mutex> lock;
    hipStream_t st = nullptr;
    size_t used = 0;
    explicit Export(const fornos_meshmap* mm) : sc(mm->sc), lock(mm->sc->xmu) {}
    // after: the solver's last queued pass (or, without one, all work queued
    // on the scene stream so far)
    fornos_status begin(size_t bytes, const fornos_solver* after)
    {
        if (!sc->xstream) HIP_TRY(hipStreamCreateWithFlags(&sc->xstream, hipStreamNonBlocking));
        if (!sc->xevent) HIP_TRY(hipEventCreateWithFlags(&sc->xevent, hipEventDisableTiming));
        st = sc->xstream;
        if (sc->xcap < bytes) {
            HIP_TRY(hipStreamSynchronize(st));
            dev_free(sc->xbuf);
            sc->xbuf = nullptr;
            sc->xcap = 0;
            if (fornos_status e = dev_alloc(&sc->xbuf, bytes)) return e;
            sc->xcap = bytes;
        }
        hipEvent_t wait = after->done_ev;
        if (!wait) {
            HIP_TRY(hipEventRecord(sc->xevent, sc->stream));
            wait = sc->xevent;
        }
        HIP_TRY(hipStreamWaitEvent(st, wait, 0));
        return FORNOS_OK;
    }
    template <typename T>
    T* take(size_t count)
    {
        T* p = (T*)(sc->xbuf + used);
        used += (sizeof(T) * std::max<size_t>(1, count) + 255) & ~(size_t)255;
        return p;
    }
    fornos_status read(void* host, const void* dev, size_t bytes)
    {
        HIP_TRY(hipGetLastError());
        if (bytes) HIP_TRY(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return FORNOS_OK;
    }
};

// workspace of any export of a map: 16 bytes per pixel at most, plus the range keys
size_t export_bytes(const fornos_meshmap* mm) { return 16 * (size_t)mm->width * mm->height + 2048; }

fornos_status export_check(const fornos_solver* s, uint32_t comps, const char* what)
{
    if (!s) return set_error(FORNOS_INVALID_ARG, std::string(what) + ": null solver");
    if (s->comps() != comps)
        return set_error(FORNOS_INVALID_ARG, std::string(what) + ": needs a solver with " + std::to_string(comps) +
                                                 " value(s) per texel");
    if (!s->mm->indices)
        return set_error(FORNOS_INVALID_ARG, std::string(what) + ": the texel set has no map indices");
    return FORNOS_OK;
}

// getMinMax + the filter + computeScaleBias (Src/image.cpp:37-47, 72-80, 199-207)
fornos_status export_range(Export& X, const fornos_solver* s, float lo, float hi, int normalize, QuantArgs& A,
                           float* minmax)
{
    const fornos_meshmap* mm = s->mm;
    hipStream_t st = X.st;
    float tmin = FLT_MAX, tmax = -FLT_MAX;
    if (mm->n) {
        uint32_t* keys = X.take<uint32_t>(2);
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, st, keys, (size_t)1, 0xffffffffu);
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, st, keys + 1, (size_t)1, 0u);
        hipLaunchKernelGGL(k_minmax, dim3(std::min<uint32_t>(1024, div_up(mm->n, kBlock))), dim3(kBlock), 0, st,
                           (const float*)s->out, mm->n, keys);
        uint32_t k[2];
        if (fornos_status e = X.read(k, keys, sizeof(k))) return e;
        auto inv = [](uint32_t key) {
            const uint32_t u = (key & 0x80000000u) ? (key & 0x7fffffffu) : ~key;
            float f;
            std::memcpy(&f, &u, 4);
            return f;
        };
        if (k[0] != 0xffffffffu) tmin = inv(k[0]);
        if (k[1] != 0u) tmax = inv(k[1]);
    }
    const bool filter = hi > lo;
    const float mn = filter ? std::fmax(tmin, lo) : tmin, mx = filter ? std::fmin(tmax, hi) : tmax;
    if (minmax) {
        minmax[0] = mn;
        minmax[1] = mx;
    }
    float scale = 1.0f, bias = 0.0f;
    if (normalize && mn != FLT_MAX && mx != -FLT_MAX) {
        scale = 1.0f / (mx - mn);
        bias = -mn * scale;
    }
    A.v = s->out;
    A.indices = mm->indices;
    A.n = mm->n;
    A.w = mm->width;
    A.h = mm->height;
    A.lo = lo;
    A.hi = hi;
    A.scale = scale;
    A.bias = bias;
    A.filter = filter ? 1 : 0;
    A.normalize = normalize ? 1 : 0;
    return FORNOS_OK;
}

// rgb (device, w*h*3) -> dilated copy back to the host
fornos_status export_finish_rgb(Export& X, const fornos_meshmap* mm, uint8_t* drgb, bool valid_from_rgb, int dilate,
                                uint8_t* rgb)
{
    hipStream_t st = X.st;
    const size_t npix = (size_t)mm->width * mm->height;
    const uint8_t* res = drgb;
    if (dilate > 0 && npix) {
        uint8_t* valid = X.take<uint8_t>(npix);
        uint8_t* out = X.take<uint8_t>(npix * 3);
        HIP_TRY(hipMemsetAsync(valid, 0, npix, st));
        if (valid_from_rgb) {
            if (mm->n)
                hipLaunchKernelGGL(k_valid_rgb, dim3(div_up(mm->n, kBlock)), dim3(kBlock), 0, st,
                                   (const uint32_t*)mm->indices, mm->n, mm->width, mm->height, (const uint8_t*)drgb,
                                   valid);
        } else if (mm->n) {
            hipLaunchKernelGGL(k_valid_texels, dim3(div_up(mm->n, kBlock)), dim3(kBlock), 0, st,
                               (const uint32_t*)mm->indices, mm->n, valid);
        }
        uint32_t* centre = X.take<uint32_t>(npix);
        HIP_TRY(hipMemcpyAsync(out, drgb, npix * 3, hipMemcpyDeviceToDevice, st));
        const dim3 grid(div_up(mm->width, kBlock), mm->height);
        hipLaunchKernelGGL(k_dilate_centre, grid, dim3(kBlock), 0, st, (const uint8_t*)valid, centre, mm->width,
                           mm->height, (uint32_t)dilate);
        hipLaunchKernelGGL(k_dilate_fill, grid, dim3(kBlock), 0, st, (const uint8_t*)drgb, out, (const uint8_t*)valid,
                           (const uint32_t*)centre, mm->width, mm->height);
        res = out;
    }
    return X.read(rgb, res, npix * 3);
}

}  // namespace

extern "C" {

fornos_status fornos_export_float_rgb8(const fornos_solver* s, float filter_lo, float filter_hi, int normalize,
                                       int dilate, uint8_t* rgb, float* minmax)
{
    if (!rgb) return set_error(FORNOS_INVALID_ARG, "fornos_export_float_rgb8: null output");
    if (fornos_status e = export_check(s, 1, "fornos_export_float_rgb8")) return e;
    const fornos_meshmap* mm = s->mm;
    Export X(mm);
    if (fornos_status e = X.begin(export_bytes(mm), s)) return e;
    QuantArgs A{};
    if (fornos_status e = export_range(X, s, filter_lo, filter_hi, normalize, A, minmax)) return e;
    const size_t npix = (size_t)mm->width * mm->height;
    uint8_t* d = X.take<uint8_t>(npix * 3);
    HIP_TRY(hipMemsetAsync(d, 0, npix * 3, X.st));
    if (mm->n) hipLaunchKernelGGL(k_quant_gray, dim3(div_up(mm->n, kBlock)), dim3(kBlock), 0, X.st, A, d);
    return export_finish_rgb(X, mm, d, false, dilate, rgb);
}

fornos_status fornos_export_float_f32(const fornos_solver* s, float filter_lo, float filter_hi, int normalize,
                                      float* image, float* minmax)
{
    if (!image) return set_error(FORNOS_INVALID_ARG, "fornos_export_float_f32: null output");
    if (fornos_status e = export_check(s, 1, "fornos_export_float_f32")) return e;
    const fornos_meshmap* mm = s->mm;
    Export X(mm);
    if (fornos_status e = X.begin(export_bytes(mm), s)) return e;
    QuantArgs A{};
    if (fornos_status e = export_range(X, s, filter_lo, filter_hi, normalize, A, minmax)) return e;
    const size_t npix = (size_t)mm->width * mm->height;
    float* d = X.take<float>(npix);
    HIP_TRY(hipMemsetAsync(d, 0, npix * sizeof(float), X.st));
    if (mm->n) hipLaunchKernelGGL(k_float_plane, dim3(div_up(mm->n, kBlock)), dim3(kBlock), 0, X.st, A, d);
    return X.read(image, d, npix * sizeof(float));
}

fornos_status fornos_export_normal_rgb8(const fornos_solver* s, int dilate, uint8_t* rgb)
{
    if (!rgb) return set_error(FORNOS_INVALID_ARG, "fornos_export_normal_rgb8: null output");
    if (fornos_status e = export_check(s, 3, "fornos_export_normal_rgb8")) return e;
    const fornos_meshmap* mm = s->mm;
    Export X(mm);
    if (fornos_status e = X.begin(export_bytes(mm), s)) return e;
    QuantArgs A{};
    A.v = s->out;
    A.indices = mm->indices;
    A.n = mm->n;
    A.w = mm->width;
    A.h = mm->height;
    const size_t npix = (size_t)mm->width * mm->height;
    uint8_t* d = X.take<uint8_t>(npix * 3);
    HIP_TRY(hipMemsetAsync(d, 0, npix * 3, X.st));
    if (mm->n) hipLaunchKernelGGL(k_quant_normal, dim3(div_up(mm->n, kBlock)), dim3(kBlock), 0, X.st, A, d);
    return export_finish_rgb(X, mm, d, true, dilate, rgb);
}

fornos_status fornos_export_vector_f32(const fornos_solver* s, float* planes)
{
    if (!planes) return set_error(FORNOS_INVALID_ARG, "fornos_export_vector_f32: null output");
    if (fornos_status e = export_check(s, 3, "fornos_export_vector_f32")) return e;
    const fornos_meshmap* mm = s->mm;
    Export X(mm);
    if (fornos_status e = X.begin(export_bytes(mm), s)) return e;
    QuantArgs A{};
    A.v = s->out;
    A.indices = mm->indices;
    A.n = mm->n;
    A.w = mm->width;
    A.h = mm->height;
    const size_t npix = (size_t)mm->width * mm->height;
    float* d = X.take<float>(npix * 3);
    HIP_TRY(hipMemsetAsync(d, 0, npix * 3 * sizeof(float), X.st));
    if (mm->n) hipLaunchKernelGGL(k_vector_planes, dim3(div_up(mm->n, kBlock)), dim3(kBlock), 0, X.st, A, d);
    return X.read(planes, d, npix * 3 * sizeof(float));
}

fornos_status fornos_solver_map_size(const fornos_solver* s, uint32_t* width, uint32_t* height)
{
    if (!s) return set_error(FORNOS_INVALID_ARG, "fornos_solver_map_size: null solver");
    if (width) *width = s->mm->width;
    if (height) *height = s->mm->height;
    return FORNOS_OK;
}

}  // extern "C"
