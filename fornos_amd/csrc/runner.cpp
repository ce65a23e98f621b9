// runner.cpp -- the bake runner: FornosRunner and its tasks
// (Src/fornos.h:129-149, Src/fornos.cpp:53-251) on top of the C ABI.
//
// start() follows the reference's assembly order: low-poly mesh and its
// normal import mode, hi-poly mesh (or the low-poly one) and its mode, the
// tangent frame when a tangent-space map is asked for, the mapping mesh of
// the chosen method, the texel set, the hi-poly BVH, the mesh mapping, then
// one task per enabled map pushed in the reference's order (thickness, bent
// normals, AO, normals, positions, height) and the mesh mapping task last,
// so that run(), which works on the back of the list, maps first.  AO and
// bent normals with the same rays are paired (fornos_solver_pair): the AO
// task's pass computes both maps.
//
// run() never waits on the device: the first call on a task submits all of
// its work (fornos_*_submit, a few launches with a completion mark after
// each), later calls poll the marks, and the call that finds the task done
// hands its finish (the export each Task::finish performs) to a host thread
// and submits the next task, so exports overlap the following computation;
// pending() holds until every finish has ended.
#include <chrono>
#include <cstring>
#include <future>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fornos_export.h"
#include "../../include/fornos_runner.h"
#include "host.h"

namespace fornos {
namespace {

// Submission parts of a task: one per 2^27 rays (or texels), 1..32.  Each
// part is its own persistent launch whose tail the next one cannot fill, so
// finer progress costs time (32 parts of an AO-256 pass on a 2048^2 map:
// +3 %; 8 parts: +0.4 %).
uint32_t parts_for(uint64_t work) { return (uint32_t)std::min<uint64_t>(32, std::max<uint64_t>(1, work >> 27)); }
constexpr uint32_t kMappingParts = 8;
constexpr uint32_t kSamplePermCount = 64 * 64;  // Src/solver_ao.cpp:34

struct MeshPtr {
    fornos_mesh* m = nullptr;
    ~MeshPtr() { fornos_mesh_destroy(m); }
};

// FornosTask (Src/fornos.h:129-137)
class Task {
public:
    virtual ~Task() {}
    virtual fornos_status submit() = 0;
    virtual fornos_status poll(float* progress, int* done) = 0;
    virtual fornos_status finish() = 0;
    virtual const char* name() const = 0;
    virtual int kind() const { return -1; }
    bool submitted = false;
    float last_progress = 0.0f;
};

// MeshMappingTask (Src/meshmapping.cpp:220-244): finish is a memory barrier
// in the reference; stream order covers it here.
class MappingTask : public Task {
public:
    explicit MappingTask(fornos_meshmap* mm) : _mm(mm) {}
    fornos_status submit() override { return fornos_meshmap_submit(_mm, kMappingParts); }
    fornos_status poll(float* p, int* d) override { return fornos_meshmap_poll(_mm, p, d); }
    fornos_status finish() override { return FORNOS_OK; }
    const char* name() const override { return "Mesh mapping"; }

private:
    fornos_meshmap* _mm;
};

// The solver tasks (Src/solver_*.cpp, XTask): finish exports the map.
class SolverTask : public Task {
public:
    SolverTask(int kind, fornos_solver* s, std::string path, int dilation, const fornos_solver_params& p,
               uint32_t texels)
        : _kind(kind), _s(s), _path(std::move(path)), _dilation(dilation), _p(p), _texels(texels)
    {
    }
    ~SolverTask() override { fornos_solver_destroy(_s); }
    fornos_status submit() override
    {
        const uint64_t work = (uint64_t)_texels * (_kind >= FORNOS_SOLVER_AO ? _p.sample_count : 1);
        return fornos_solver_submit(_s, parts_for(work));
    }
    fornos_status poll(float* p, int* d) override { return fornos_solver_poll(_s, p, d); }
    fornos_status finish() override
    {
        if (_path.empty()) return FORNOS_OK;
        const char* path = _path.c_str();
        switch (_kind) {
        case FORNOS_SOLVER_AO:         // Src/solver_ao.cpp:155-163
        case FORNOS_SOLVER_THICKNESS:  // Src/solver_thickness.cpp:156-164
            exported = true;
            return fornos_export_float_image(_s, path, 0.0f, 0.0f, 1, _dilation, range);
        case FORNOS_SOLVER_HEIGHT:     // Src/solver_height.cpp:103-119
            exported = true;
            return fornos_export_float_image(_s, path, 0.0f, _p.max_distance, _p.normalize_output, _dilation, range);
        case FORNOS_SOLVER_NORMALS:       // Src/solver_normals.cpp:116-123
        case FORNOS_SOLVER_BENT_NORMALS:  // Src/solver_bentnormals.cpp:166-174
            return fornos_export_normal_image(_s, path, _dilation);
        default:                          // Src/solver_position.cpp:104-111
            return fornos_export_vector_image(_s, path);
        }
    }
    const char* name() const override
    {
        switch (_kind) {
        case FORNOS_SOLVER_HEIGHT: return "Height";
        case FORNOS_SOLVER_POSITION: return "Position";
        case FORNOS_SOLVER_NORMALS: return "Normals";
        case FORNOS_SOLVER_AO: return "Ambient Occlusion";
        case FORNOS_SOLVER_BENT_NORMALS: return "Bent normals";
        default: return "Thickness";
        }
    }
    int kind() const override { return _kind; }
    fornos_solver* solver() const { return _s; }
    bool exported = false;
    float range[2] = {0.0f, 0.0f};

private:
    int _kind;
    fornos_solver* _s;
    std::string _path;
    int _dilation;
    fornos_solver_params _p;
    uint32_t _texels;
};

bool has_text(const char* s) { return s && s[0] != '\0'; }

fornos_status apply_normals(fornos_mesh* m, int mode)
{
    switch (mode) {
    case FORNOS_NORMALS_IMPORT: return FORNOS_OK;
    case FORNOS_NORMALS_PER_FACE: return fornos_mesh_compute_face_normals(m);
    case FORNOS_NORMALS_PER_VERTEX: return fornos_mesh_compute_vertex_normals(m);
    }
    return set_error(FORNOS_INVALID_ARG, "unknown normal import mode " + std::to_string(mode));
}

}  // namespace
}  // namespace fornos

using namespace fornos;

struct fornos_runner {
    std::unique_ptr<MeshPtr> lo, hi, dirs;
    fornos_scene* scene = nullptr;
    fornos_meshmap* mm = nullptr;
    std::vector<std::unique_ptr<Task>> tasks;  // run from the back
    // finished tasks whose finish (the export) runs on a host thread while
    // the next task computes; with its status and message
    struct Finishing {
        std::unique_ptr<Task> task;
        std::future<std::pair<fornos_status, std::string>> result;
    };
    std::vector<Finishing> finishing;
    const char* last = nullptr;  // name of the task handed to `finishing` last
    std::vector<std::pair<int, std::pair<float, float>>> ranges;  // exported float maps
    // Collects the finishes that have ended (wait: all of them); the first
    // failure is returned with its message.
    fornos_status collect(bool wait)
    {
        fornos_status first = FORNOS_OK;
        std::string msg;
        for (auto it = finishing.begin(); it != finishing.end();) {
            if (!wait && it->result.wait_for(std::chrono::seconds(0)) != std::future_status::ready) {
                ++it;
                continue;
            }
            const auto res = it->result.get();
            if (auto* s = dynamic_cast<SolverTask*>(it->task.get()))
                if (s->exported && res.first == FORNOS_OK) ranges.push_back({s->kind(), {s->range[0], s->range[1]}});
            if (res.first != FORNOS_OK && first == FORNOS_OK) {
                first = res.first;
                msg = res.second;
            }
            it = finishing.erase(it);
        }
        return first == FORNOS_OK ? FORNOS_OK : set_error(first, msg);
    }
    void clear()
    {
        (void)collect(true);  // exports still reading results end first
        last = nullptr;
        tasks.clear();  // solvers before the mapping and the scene they use
        fornos_meshmap_destroy(mm);
        mm = nullptr;
        fornos_scene_destroy(scene);
        scene = nullptr;
        lo.reset();
        hi.reset();
        dirs.reset();
        ranges.clear();
    }
    ~fornos_runner() { clear(); }
};

extern "C" {

void fornos_parameters_defaults(fornos_parameters* p)
{
    if (!p) return;
    *p = fornos_parameters{};
    fornos_params_shared& s = p->shared;
    s.lo_poly_mesh_path = "";
    s.hi_poly_mesh_path = "";
    s.lo_poly_mesh_normal = FORNOS_NORMALS_IMPORT;
    s.hi_poly_mesh_normal = FORNOS_NORMALS_IMPORT;
    s.bvh_tris_per_node = 8;
    s.tex_width = 2048;
    s.tex_height = 2048;
    s.tex_dilation = 16;
    s.ignore_backfaces = 1;
    s.mapping = FORNOS_MAPPING_SMOOTH;
    s.mapping_edge = 0.05f;
    s.sample_seed = 1234;
    fornos_params_map* maps[6] = {&p->height, &p->positions, &p->normals, &p->ao, &p->bent_normals, &p->thickness};
    for (fornos_params_map* m : maps) {
        m->enabled = 0;
        m->output_path = "";
        m->sample_count = 256;
        m->min_distance = 0.01f;
        m->max_distance = 10.0f;
        m->tangent_space = 1;
        m->normalize_output = 1;
    }
    p->height.max_distance = 0.0f;  // FornosParameters_SolverHeight::maxDistance
}

fornos_status fornos_runner_create(fornos_runner** out)
{
    if (!out) return set_error(FORNOS_INVALID_ARG, "fornos_runner_create: null argument");
    *out = new (std::nothrow) fornos_runner();
    return *out ? FORNOS_OK : set_error(FORNOS_OOM, "fornos_runner_create: out of host memory");
}

// Src/fornos.cpp:53-236
fornos_status fornos_runner_start(fornos_runner* r, const fornos_parameters* p)
{
    if (!r || !p) return set_error(FORNOS_INVALID_ARG, "fornos_runner_start: null argument");
    r->clear();
    const fornos_params_shared& sh = p->shared;
    if (sh.tex_width < 1 || sh.tex_height < 1 || sh.tex_width > 16384 || sh.tex_height > 16384)
        return set_error(FORNOS_INVALID_ARG, "texture size must be in [1, 16384]");
    if (sh.bvh_tris_per_node < 1) return set_error(FORNOS_INVALID_ARG, "bvhTrisPerNode must be >= 1");
    if (sh.tex_dilation < 0) return set_error(FORNOS_INVALID_ARG, "texDilation must be >= 0");
    if (sh.mapping < FORNOS_MAPPING_SMOOTH || sh.mapping > FORNOS_MAPPING_HYBRID)
        return set_error(FORNOS_INVALID_ARG, "unknown mesh mapping method " + std::to_string(sh.mapping));
    if (sh.mapping == FORNOS_MAPPING_HYBRID && !(sh.mapping_edge > 0.0f))
        return set_error(FORNOS_INVALID_ARG, "the Hybrid mapping needs mappingEdge > 0");

    // low-poly mesh and its normals
    r->lo.reset(new MeshPtr());
    if (!has_text(sh.lo_poly_mesh_path) || fornos_mesh_load(sh.lo_poly_mesh_path, &r->lo->m) != FORNOS_OK)
        return set_error(FORNOS_INVALID_ARG, std::string("Missing low poly mesh") +
                                                 (has_text(sh.lo_poly_mesh_path) ? std::string(" (") + fornos_last_error() + ")" : ""));
    if (fornos_status st = apply_normals(r->lo->m, sh.lo_poly_mesh_normal)) return st;

    // hi-poly mesh: the low-poly one when no path is given
    fornos_mesh* hi = r->lo->m;
    if (has_text(sh.hi_poly_mesh_path)) {
        r->hi.reset(new MeshPtr());
        if (fornos_mesh_load(sh.hi_poly_mesh_path, &r->hi->m) != FORNOS_OK)
            return set_error(FORNOS_INVALID_ARG, std::string("Missing high poly mesh (") + fornos_last_error() + ")");
        if (fornos_status st = apply_normals(r->hi->m, sh.hi_poly_mesh_normal)) return st;
        hi = r->hi->m;
    }

    const bool tangent = (p->normals.enabled && p->normals.tangent_space) ||
                         (p->bent_normals.enabled && p->bent_normals.tangent_space);
    if (tangent)
        if (fornos_status st = fornos_mesh_compute_tangent_space(r->lo->m)) return st;

    // mapping mesh of the method (Src/fornos.cpp:96-148)
    fornos_mesh* dirs = nullptr;
    if (sh.mapping != FORNOS_MAPPING_LOWPOLY_NORMALS) {
        if (sh.lo_poly_mesh_normal != FORNOS_NORMALS_PER_VERTEX) {
            r->dirs.reset(new MeshPtr());
            if (fornos_status st = fornos_mesh_copy(r->lo->m, &r->dirs->m)) return st;
            if (fornos_status st = fornos_mesh_compute_vertex_normals_aggressive(r->dirs->m)) return st;
            dirs = r->dirs->m;
        } else {
            dirs = r->lo->m;
        }
    }

    // hi-poly BVH, texel set and mesh mapping on the device
    fornos_mesh_view hv, lv, dv;
    if (fornos_status st = fornos_mesh_view_get(hi, &hv)) return st;
    if (fornos_status st = fornos_mesh_view_get(r->lo->m, &lv)) return st;
    if (dirs)
        if (fornos_status st = fornos_mesh_view_get(dirs, &dv)) return st;
    fornos_bvh_params bp{FORNOS_BVH_REFERENCE, (uint32_t)sh.bvh_tris_per_node, 8192};
    if (fornos_status st = fornos_scene_create(&hv, &bp, &r->scene)) return st;
    const float edge = sh.mapping == FORNOS_MAPPING_HYBRID ? sh.mapping_edge : 0.0f;
    if (fornos_status st = fornos_meshmap_create_from_mesh_hybrid(r->scene, &lv, dirs ? &dv : nullptr,
                                                                  (uint32_t)sh.tex_width, (uint32_t)sh.tex_height,
                                                                  edge, sh.ignore_backfaces ? 1 : 0, &r->mm))
        return st;

    // one task per enabled map, in the reference's push order
    struct Spec {
        const fornos_params_map* m;
        int kind;
    };
    const Spec specs[6] = {{&p->thickness, FORNOS_SOLVER_THICKNESS}, {&p->bent_normals, FORNOS_SOLVER_BENT_NORMALS},
                           {&p->ao, FORNOS_SOLVER_AO},               {&p->normals, FORNOS_SOLVER_NORMALS},
                           {&p->positions, FORNOS_SOLVER_POSITION},  {&p->height, FORNOS_SOLVER_HEIGHT}};
    std::vector<float> table;
    std::vector<uint32_t> rnd;
    uint32_t texels = 0;
    if (fornos_status st = fornos_meshmap_texel_count(r->mm, &texels)) return st;
    for (const Spec& sp : specs) {
        const fornos_params_map& m = *sp.m;
        if (!m.enabled) continue;
        const bool hemi = sp.kind >= FORNOS_SOLVER_AO;
        fornos_solver_params sprm{};
        sprm.tangent_space = m.tangent_space ? 1 : 0;
        sprm.normalize_output = m.normalize_output ? 1 : 0;
        sprm.max_distance = m.max_distance;
        if (hemi) {
            if (m.sample_count < 1) return set_error(FORNOS_INVALID_ARG, "sampleCount must be >= 1");
            sprm.sample_count = (uint32_t)m.sample_count;
            sprm.min_distance = m.min_distance;
            table.assign((size_t)kSamplePermCount * sprm.sample_count * 3, 0.0f);
            rnd.assign((size_t)kSamplePermCount * 2, 0u);
            if (fornos_status st = fornos_sample_table(sprm.sample_count, kSamplePermCount, sh.sample_seed,
                                                       table.data(), rnd.data()))
                return st;
        }
        fornos_solver* s = nullptr;
        if (fornos_status st = fornos_solver_create(sp.kind, &sprm, r->mm, hemi ? table.data() : nullptr,
                                                    hemi ? kSamplePermCount : 0, &s))
            return st;
        const int dilation = sp.kind == FORNOS_SOLVER_POSITION ? 0 : sh.tex_dilation;
        r->tasks.emplace_back(
            new SolverTask(sp.kind, s, has_text(m.output_path) ? m.output_path : "", dilation, sprm, texels));
    }
    // AO and bent normals casting the same rays share one traversal: the AO
    // task (run first) computes both, the BN task finds its work done
    auto same_bits = [](float a, float b) { return std::memcmp(&a, &b, sizeof(float)) == 0; };
    if (p->ao.enabled && p->bent_normals.enabled && p->ao.sample_count == p->bent_normals.sample_count &&
        same_bits(p->ao.min_distance, p->bent_normals.min_distance) &&
        same_bits(p->ao.max_distance, p->bent_normals.max_distance)) {
        fornos_solver *ao = nullptr, *bn = nullptr;
        for (auto& t : r->tasks) {
            if (t->kind() == FORNOS_SOLVER_AO) ao = static_cast<SolverTask*>(t.get())->solver();
            if (t->kind() == FORNOS_SOLVER_BENT_NORMALS) bn = static_cast<SolverTask*>(t.get())->solver();
        }
        if (fornos_status st = fornos_solver_pair(ao, bn)) return st;
    }
    r->tasks.emplace_back(new MappingTask(r->mm));
    return FORNOS_OK;
}

int fornos_runner_pending(const fornos_runner* r) { return r && (!r->tasks.empty() || !r->finishing.empty()); }

// Src/fornos.cpp:239-251
fornos_status fornos_runner_run(fornos_runner* r)
{
    if (!r) return set_error(FORNOS_INVALID_ARG, "fornos_runner_run: null handle");
    if (fornos_status st = r->collect(false)) return st;
    if (r->tasks.empty()) return FORNOS_OK;
    Task& t = *r->tasks.back();
    if (!t.submitted) {
        t.submitted = true;
        return t.submit();
    }
    int done = 0;
    if (fornos_status st = t.poll(&t.last_progress, &done)) return st;
    if (!done) return FORNOS_OK;
    // the task's finish runs on a host thread (exports wait on the device for
    // the task's own work only), and the next task starts at once
    std::unique_ptr<Task> task = std::move(r->tasks.back());
    r->tasks.pop_back();
    Task* raw = task.get();
    r->last = raw->name();
    r->finishing.push_back({std::move(task), std::async(std::launch::async, [raw] {
                                const fornos_status st = raw->finish();
                                return std::make_pair(st, std::string(st ? fornos_last_error() : ""));
                            })});
    if (r->tasks.empty()) return FORNOS_OK;
    Task& next = *r->tasks.back();
    next.submitted = true;
    return next.submit();
}

fornos_status fornos_runner_run_all(fornos_runner* r)
{
    if (!r) return set_error(FORNOS_INVALID_ARG, "fornos_runner_run_all: null handle");
    while (!r->tasks.empty()) {
        const size_t before = r->tasks.size();
        const bool was_submitted = r->tasks.back()->submitted;
        if (fornos_status st = fornos_runner_run(r)) return st;
        if (r->tasks.size() == before && was_submitted) std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    return r->collect(true);
}

fornos_status fornos_runner_current_task(fornos_runner* r, const char** name, float* progress)
{
    if (!r) return set_error(FORNOS_INVALID_ARG, "fornos_runner_current_task: null handle");
    if (r->tasks.empty()) {
        // while exports still run, the last task handed over stays the current one
        if (name) *name = r->finishing.empty() ? nullptr : r->last;
        if (progress) *progress = 1.0f;
        return FORNOS_OK;
    }
    Task& t = *r->tasks.back();
    if (t.submitted) {
        int done = 0;
        if (fornos_status st = t.poll(&t.last_progress, &done)) return st;
    }
    if (name) *name = t.name();
    if (progress) *progress = t.last_progress;
    return FORNOS_OK;
}

int fornos_runner_exported_range(const fornos_runner* r, int task, float* minmax)
{
    if (!r) return 0;
    for (const auto& e : r->ranges)
        if (e.first == task) {
            if (minmax) {
                minmax[0] = e.second.first;
                minmax[1] = e.second.second;
            }
            return 1;
        }
    return 0;
}

void fornos_runner_destroy(fornos_runner* r) { delete r; }

}  // extern "C"
