// shim_bake.cpp -- a C++ host of the reference-signature classes
// (fornos_amd/csrc/fornos.hpp): the sphere workload (UV sphere 32 x 16 over a
// 6-times subdivided displaced icosphere, 512^2 map, culling mapping) baked
// through MeshMapping::init / runStep / progress (Src/meshmapping.h:42-57) and
// each solver's init / runStep / getResults / progress (Src/solver_ao.h:36-96
// and analogues), one runStep per loop iteration as FornosRunner::run does
// (Src/fornos.cpp:239-251).  getResults() returns the reference's new[] array
// of workCount entries, which this program delete[]s.
//
//     shim_bake OUT_DIR
//
// writes the mapping (coords, tidx) and every solver's getResults() array as
// raw little-endian files, and meta.txt with the counts; the GPU test compares
// them with the C-ABI path.
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "../../fornos_amd/csrc/fornos.hpp"

namespace {

void write(const std::string& path, const void* data, size_t bytes)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(data, 1, bytes, f) != bytes) throw std::runtime_error("cannot write " + path);
    std::fclose(f);
}

struct MeshDel {
    void operator()(fornos_mesh* m) const { fornos_mesh_destroy(m); }
};
struct TexelsDel {
    void operator()(const fornos_texels* t) const { fornos_texels_destroy(const_cast<fornos_texels*>(t)); }
};

// runStep until done; returns the number of steps, checks progress is monotone and ends at 1
template <typename T>
int run_steps(T& x, const char* what)
{
    int steps = 0;
    float last = x.progress();
    bool done = false;
    while (!done) {
        done = x.runStep();
        ++steps;
        const float p = x.progress();
        if (p < last) throw std::runtime_error(std::string(what) + ": progress went back");
        last = p;
    }
    if (last != 1.0f) throw std::runtime_error(std::string(what) + ": progress did not reach 1");
    return steps;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: shim_bake OUT_DIR\n");
        return 2;
    }
    const std::string out = argv[1];
    try {
        fornos_mesh *lo = nullptr, *hi = nullptr, *dirs = nullptr;
        const fornos_gen_params plo{1.0f, 32, 16}, phi{0.02f, 6, 0};
        fornos::check(fornos_mesh_generate(FORNOS_GEN_UV_SPHERE, 0, &plo, &lo), "low-poly mesh");
        std::unique_ptr<fornos_mesh, MeshDel> lo_(lo);
        fornos::check(fornos_mesh_generate(FORNOS_GEN_ICOSPHERE_WAVES, 7, &phi, &hi), "hi-poly mesh");
        std::unique_ptr<fornos_mesh, MeshDel> hi_(hi);
        // the Smooth mapping method: ray directions from aggressive vertex
        // normals of a copy of the low-poly mesh (Src/fornos.cpp:100-118)
        fornos::check(fornos_mesh_copy(lo, &dirs), "mapping mesh");
        std::unique_ptr<fornos_mesh, MeshDel> dirs_(dirs);
        fornos::check(fornos_mesh_compute_vertex_normals_aggressive(dirs), "mapping normals");
        fornos_texels* tx = nullptr;
        fornos::check(fornos_texels_create(lo, dirs, 512, 512, &tx), "CompressedMapUV");
        std::shared_ptr<const fornos_texels> map(tx, TexelsDel());

        auto scene = std::make_shared<fornos::Scene>(hi);
        auto mm = std::make_shared<fornos::MeshMapping>();
        mm->init(map, scene, true);
        const int map_steps = run_steps(*mm, "MeshMapping");
        const uint32_t n = mm->texelCount();
        const size_t work = ((size_t)n + 63) / 64 * 64;
        std::vector<float> coords(4 * (size_t)n);
        std::vector<uint32_t> tidx(n);
        fornos::check(fornos_meshmap_read(mm->handle(), coords.data(), tidx.data()), "mapping readback");
        write(out + "/coords.bin", coords.data(), coords.size() * 4);
        write(out + "/tidx.bin", tidx.data(), tidx.size() * 4);

        std::string meta = "texels " + std::to_string(n) + "\nwork " + std::to_string(work) + "\nmapping_steps " +
                           std::to_string(map_steps) + "\n";
        auto bake = [&](fornos::SolverBase& s, float* (*get)(fornos::SolverBase&), int comps, const char* name) {
            const int steps = run_steps(s, name);
            float* r = get(s);
            write(out + "/" + name + ".bin", r, sizeof(float) * work * comps);
            delete[] r;
            meta += std::string(name) + "_steps " + std::to_string(steps) + "\n";
        };
        fornos::AmbientOcclusionSolver ao({32, 0.0f, 0.5f});
        ao.init(map, mm);
        bake(ao, [](fornos::SolverBase& s) { return static_cast<fornos::AmbientOcclusionSolver&>(s).getResults(); }, 1, "ao");
        fornos::BentNormalsSolver bn({32, 0.0f, 0.5f, false});
        bn.init(map, mm);
        bake(bn, [](fornos::SolverBase& s) { return static_cast<fornos::BentNormalsSolver&>(s).getResults(); }, 3,
             "bentnormals");
        fornos::ThicknessSolver th({16, 0.0f, 0.5f});
        th.init(map, mm);
        bake(th, [](fornos::SolverBase& s) { return static_cast<fornos::ThicknessSolver&>(s).getResults(); }, 1,
             "thickness");
        fornos::HeightSolver hs({true, 0.0f});
        hs.init(map, mm);
        bake(hs, [](fornos::SolverBase& s) { return static_cast<fornos::HeightSolver&>(s).getResults(); }, 1, "height");
        fornos::NormalsSolver ns({false});
        ns.init(map, mm);
        bake(ns, [](fornos::SolverBase& s) { return static_cast<fornos::NormalsSolver&>(s).getResults(); }, 3, "normals");
        fornos::PositionSolver ps;
        ps.init(map, mm);
        bake(ps, [](fornos::SolverBase& s) { return static_cast<fornos::PositionSolver&>(s).getResults(); }, 3,
             "position");
        write(out + "/meta.txt", meta.data(), meta.size());
    } catch (const fornos::Error& e) {
        std::fprintf(stderr, "shim_bake: %s (status %d)\n", e.what(), (int)e.status);
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "shim_bake: %s\n", e.what());
        return 1;
    }
    return 0;
}
