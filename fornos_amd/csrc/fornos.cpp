// fornos.cpp -- reference-interface C++ classes over the C ABI.
#include "fornos.hpp"

#include <cstring>

namespace fornos {

void check(fornos_status st, const char* where)
{
    if (st != FORNOS_OK) throw Error(st, std::string(where) + ": " + fornos_last_error());
}

Scene::Scene(const fornos_mesh* hiPoly, uint32_t bvhTrisPerNode, uint32_t maxDepth)
{
    fornos_mesh_view v;
    check(fornos_mesh_view_get(hiPoly, &v), "Scene");
    fornos_bvh_params p{FORNOS_BVH_REFERENCE, bvhTrisPerNode, maxDepth};
    check(fornos_scene_create(&v, &p, &_scene), "Scene");
}

Scene::~Scene() { fornos_scene_destroy(_scene); }

MeshMapping::~MeshMapping() { fornos_meshmap_destroy(_mm); }

// Src/meshmapping.cpp:128-184
void MeshMapping::init(std::shared_ptr<const fornos_texels> map, std::shared_ptr<Scene> scene, bool cullBackfaces)
{
    _map = map;
    _scene = scene;
    fornos_texel_view tv;
    check(fornos_texels_view_get(map.get(), &tv), "MeshMapping::init");
    fornos_meshmap_destroy(_mm);
    _mm = nullptr;
    check(fornos_meshmap_create(scene->handle(), &tv, cullBackfaces ? 1 : 0, &_mm), "MeshMapping::init");
}

// Src/meshmapping.cpp:186-218
bool MeshMapping::runStep()
{
    int done = 0;
    check(fornos_meshmap_run_step(_mm, &done), "MeshMapping::runStep");
    return done != 0;
}

float MeshMapping::progress() const
{
    float p = 0;
    check(fornos_meshmap_progress(_mm, &p), "MeshMapping::progress");
    return p;
}

uint32_t MeshMapping::texelCount() const
{
    uint32_t n = 0;
    check(fornos_meshmap_texel_count(_mm, &n), "MeshMapping::texelCount");
    return n;
}

SolverBase::~SolverBase() { fornos_solver_destroy(_s); }

void SolverBase::initSolver(std::shared_ptr<MeshMapping> mm, const float* table, uint32_t nPerm)
{
    _mm = mm;
    fornos_solver_destroy(_s);
    _s = nullptr;
    check(fornos_solver_create(_kind, &_p, mm->handle(), table, nPerm, &_s), "Solver::init");
}

bool SolverBase::runStep()
{
    int done = 0;
    check(fornos_solver_run_step(_s, &done), "Solver::runStep");
    return done != 0;
}

float SolverBase::progress() const
{
    float p = 0;
    check(fornos_solver_progress(_s, &p), "Solver::progress");
    return p;
}

// getResults(): a new[] array of workCount entries (Src/solver_ao.cpp:133-138,
// Src/solver_height.cpp:76-84); entries past the texel count are zero.
float* SolverBase::readPadded(uint32_t comps) const
{
    const uint32_t n = _mm->texelCount();
    const size_t work = ((size_t)n + 63) / 64 * 64;
    float* r = new float[work * comps];
    std::memset(r, 0, sizeof(float) * work * comps);
    check(fornos_solver_read(_s, r, sizeof(float) * (size_t)n * comps), "Solver::getResults");
    return r;
}

static std::vector<float> samples(size_t S, uint32_t seed)
{
    std::vector<float> t(S * kSamplePermCount * 3);
    check(fornos_sample_table((uint32_t)S, kSamplePermCount, seed, t.data(), nullptr), "computeSamples");
    return t;
}

AmbientOcclusionSolver::AmbientOcclusionSolver(const Params& p, uint32_t seed)
    : SolverBase(FORNOS_SOLVER_AO, fornos_solver_params{(uint32_t)p.sampleCount, p.minDistance, p.maxDistance, 0, 0}), _seed(seed) {}
void AmbientOcclusionSolver::init(std::shared_ptr<const fornos_texels>, std::shared_ptr<MeshMapping> mm)
{
    _table = samples(_p.sample_count, _seed);
    initSolver(mm, _table.data(), kSamplePermCount);
}
float* AmbientOcclusionSolver::getResults() { return readPadded(1); }

BentNormalsSolver::BentNormalsSolver(const Params& p, uint32_t seed)
    : SolverBase(FORNOS_SOLVER_BENT_NORMALS, fornos_solver_params{(uint32_t)p.sampleCount, p.minDistance, p.maxDistance, p.tangentSpace ? 1 : 0, 0}), _seed(seed) {}
void BentNormalsSolver::init(std::shared_ptr<const fornos_texels>, std::shared_ptr<MeshMapping> mm)
{
    _table = samples(_p.sample_count, _seed);
    initSolver(mm, _table.data(), kSamplePermCount);
}
float* BentNormalsSolver::getResults() { return readPadded(3); }

ThicknessSolver::ThicknessSolver(const Params& p, uint32_t seed)
    : SolverBase(FORNOS_SOLVER_THICKNESS, fornos_solver_params{(uint32_t)p.sampleCount, p.minDistance, p.maxDistance, 0, 0}), _seed(seed) {}
void ThicknessSolver::init(std::shared_ptr<const fornos_texels>, std::shared_ptr<MeshMapping> mm)
{
    _table = samples(_p.sample_count, _seed);
    initSolver(mm, _table.data(), kSamplePermCount);
}
float* ThicknessSolver::getResults() { return readPadded(1); }

HeightSolver::HeightSolver(const Params& p)
    : SolverBase(FORNOS_SOLVER_HEIGHT, fornos_solver_params{0, 0.0f, p.maxDistance, 0, p.normalizeOutput ? 1 : 0}) {}
void HeightSolver::init(std::shared_ptr<const fornos_texels>, std::shared_ptr<MeshMapping> mm) { initSolver(mm, nullptr, 0); }
float* HeightSolver::getResults() { return readPadded(1); }

NormalsSolver::NormalsSolver(const Params& p)
    : SolverBase(FORNOS_SOLVER_NORMALS, fornos_solver_params{0, 0.0f, 0.0f, p.tangentSpace ? 1 : 0, 0}) {}
void NormalsSolver::init(std::shared_ptr<const fornos_texels>, std::shared_ptr<MeshMapping> mm) { initSolver(mm, nullptr, 0); }
float* NormalsSolver::getResults() { return readPadded(3); }

PositionSolver::PositionSolver() : SolverBase(FORNOS_SOLVER_POSITION, fornos_solver_params{0, 0.0f, 0.0f, 0, 0}) {}
void PositionSolver::init(std::shared_ptr<const fornos_texels>, std::shared_ptr<MeshMapping> mm) { initSolver(mm, nullptr, 0); }
float* PositionSolver::getResults() { return readPadded(3); }

}  // namespace fornos
