// fornos.hpp -- the reference's host interface for the baking path, kept
// method for method, on top of the C ABI (include/fornos_hip.h).
//
//   MeshMapping            Src/meshmapping.h:42-75   init / runStep / progress
//   AmbientOcclusionSolver Src/solver_ao.h:36-96     init / runStep / getResults / progress
//   BentNormalsSolver      Src/solver_bentnormals.h
//   ThicknessSolver        Src/solver_thickness.h
//   HeightSolver, NormalsSolver, PositionSolver      Src/solver_{height,normals,position}.h
//
// The GL buffers of the reference become device memory owned by the C-ABI
// handles; getResults() keeps the reference's contract of returning a new[]
// array of workCount (texel count rounded up to 64) entries that the caller
// delete[]s.  Errors that the reference logged or left undefined throw
// fornos::Error.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fornos_hip.h"
#include "../../include/fornos_host.h"

namespace fornos {

struct Error : std::runtime_error {
    fornos_status status;
    Error(fornos_status st, const std::string& what) : std::runtime_error(what), status(st) {}
};
void check(fornos_status st, const char* where);

// The hi-poly mesh with its BVH, resident on the device (the reference keeps
// Mesh + BVH on the host and uploads them in MeshMapping::init).
class Scene {
public:
    Scene(const fornos_mesh* hiPoly, uint32_t bvhTrisPerNode = 8, uint32_t maxDepth = 8192);
    ~Scene();
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;
    fornos_scene* handle() const { return _scene; }

private:
    fornos_scene* _scene = nullptr;
};

class MeshMapping {
public:
    ~MeshMapping();
    void init(std::shared_ptr<const fornos_texels> map, std::shared_ptr<Scene> scene, bool cullBackfaces = false);
    bool runStep();
    float progress() const;
    fornos_meshmap* handle() const { return _mm; }
    uint32_t texelCount() const;

private:
    std::shared_ptr<const fornos_texels> _map;
    std::shared_ptr<Scene> _scene;
    fornos_meshmap* _mm = nullptr;
};

// One class per solver kind, all sharing this implementation.
class SolverBase {
public:
    virtual ~SolverBase();
    bool runStep();
    float progress() const;
    fornos_solver* handle() const { return _s; }

protected:
    SolverBase(int kind, const fornos_solver_params& p) : _kind(kind), _p(p) {}
    void initSolver(std::shared_ptr<MeshMapping> mm, const float* table, uint32_t nPerm);
    float* readPadded(uint32_t comps) const;
    int _kind;
    fornos_solver_params _p;
    std::shared_ptr<MeshMapping> _mm;
    std::vector<float> _table;
    fornos_solver* _s = nullptr;
};

// Sample table of the hemisphere solvers: the reference draws it per solver
// from std::random_device (Src/solver_ao.cpp:38-44); here it is seeded.
constexpr uint32_t kSamplePermCount = 64 * 64;  // Src/solver_ao.cpp:34

class AmbientOcclusionSolver : public SolverBase {
public:
    struct Params { size_t sampleCount; float minDistance; float maxDistance; };
    explicit AmbientOcclusionSolver(const Params& p, uint32_t seed = 1234);
    void init(std::shared_ptr<const fornos_texels> map, std::shared_ptr<MeshMapping> meshMapping);
    float* getResults();  // workCount floats, new[]
private:
    uint32_t _seed;
};

class BentNormalsSolver : public SolverBase {
public:
    struct Params { size_t sampleCount; float minDistance; float maxDistance; bool tangentSpace; };
    explicit BentNormalsSolver(const Params& p, uint32_t seed = 1234);
    void init(std::shared_ptr<const fornos_texels> map, std::shared_ptr<MeshMapping> meshMapping);
    float* getResults();  // workCount * 3 floats (Vector3), new[]
private:
    uint32_t _seed;
};

class ThicknessSolver : public SolverBase {
public:
    struct Params { size_t sampleCount; float minDistance; float maxDistance; };
    explicit ThicknessSolver(const Params& p, uint32_t seed = 1234);
    void init(std::shared_ptr<const fornos_texels> map, std::shared_ptr<MeshMapping> meshMapping);
    float* getResults();
private:
    uint32_t _seed;
};

class HeightSolver : public SolverBase {
public:
    struct Params { bool normalizeOutput; float maxDistance; };
    explicit HeightSolver(const Params& p);
    void init(std::shared_ptr<const fornos_texels> map, std::shared_ptr<MeshMapping> meshMapping);
    float* getResults();
};

class NormalsSolver : public SolverBase {
public:
    struct Params { bool tangentSpace; };
    explicit NormalsSolver(const Params& p);
    void init(std::shared_ptr<const fornos_texels> map, std::shared_ptr<MeshMapping> meshMapping);
    float* getResults();
};

class PositionSolver : public SolverBase {
public:
    PositionSolver();
    void init(std::shared_ptr<const fornos_texels> map, std::shared_ptr<MeshMapping> meshMapping);
    float* getResults();
};

}  // namespace fornos
