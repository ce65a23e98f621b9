// host.h -- internal host-side types of libfornos_hip (not part of the ABI).
#pragma once
#include <cstdint>
#include <string>
#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/fornos_host.h"
#include "../../include/fornos_hip.h"
#include "vec.h"

// The reference's Mesh (Src/mesh.h:36-77): shared attribute arrays plus one
// {position, texcoord, normal} index triple per vertex and three vertex
// indices per triangle.
struct fornos_mesh {
    std::vector<fornos::V3> positions;
    std::vector<fornos::V2> texcoords;
    std::vector<fornos::V3> normals;
    std::vector<fornos::V3> tangents;    // per vertex (may be empty)
    std::vector<fornos::V3> bitangents;  // per vertex (may be empty)
    std::vector<uint32_t> vertices;      // 3 per vertex
    std::vector<uint32_t> triangles;     // 3 per triangle

    uint32_t nv() const { return (uint32_t)(vertices.size() / 3); }
    uint32_t ntri() const { return (uint32_t)(triangles.size() / 3); }
    fornos::V3 pos_of(uint32_t vtx) const { return positions[vertices[3 * vtx]]; }
};

// MapUV (Src/compute.h:214-241) and CompressedMapUV (:245-259) together.
struct fornos_texels {
    uint32_t width = 0, height = 0;
    std::vector<float> map_pos, map_dir, map_nrm, map_tan, map_bit;  // W*H*3 each
    std::vector<float> pos, dir, nrm, tan, bit;                      // count*3 each
    std::vector<uint32_t> indices;
};

namespace fornos {

fornos_status set_error(fornos_status st, const std::string& msg);
fornos_status ok();

// Runs fn(begin, end) over [0, n) split into contiguous chunks of at least
// `grain` items on up to 16 host threads (the calling thread takes the first).
template <typename Fn>
void parallel_chunks(size_t n, Fn fn, size_t grain = 64)
{
    const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(hw, 16), std::max<size_t>(1, n / std::max<size_t>(1, grain)));
    if (nt <= 1) {
        fn((size_t)0, n);
        return;
    }
    std::vector<std::thread> pool;
    const size_t per = (n + nt - 1) / nt;
    for (size_t t = 1; t < nt; ++t) {
        const size_t b = t * per, e = std::min(n, b + per);
        if (b < e) pool.emplace_back([=] { fn(b, e); });
    }
    fn((size_t)0, std::min(n, per));
    for (std::thread& th : pool) th.join();
}

// mesh.cpp
void vertex_normals(fornos_mesh& m);             // Src/mesh.cpp:650-678
void vertex_normals_aggressive(fornos_mesh& m);  // Src/mesh.cpp:680-721
void tangent_space(fornos_mesh& m);              // Src/mesh.cpp:724-768
void face_normals(fornos_mesh& m);               // Src/mesh.cpp:627-648 (meshio.cpp)

// texels.cpp
// nullptr on success, else the reason
const char* build_texels(const fornos_mesh& lo, const fornos_mesh* dirs, uint32_t W, uint32_t H,
                         fornos_texels& out, float edge = 0.0f);

// Reference-layout BVH, flattened (bvh.cpp).
struct FlatBVH {
    // 8 words per node: min.xyz, first soup triangle, max.xyz, jump | leaf flag + count
    std::vector<uint32_t> nodes;
    std::vector<uint32_t> soup_tri;  // mesh triangle id per soup triangle
    uint32_t n_nodes = 0, n_leaves = 0, max_leaf = 0, depth = 0;
};
constexpr uint32_t kLeafFlag = 0x80000000u;
void build_reference_bvh(const fornos_mesh& m, uint32_t max_tris, uint32_t max_depth, FlatBVH& out);
void export_reference_nodes(const FlatBVH& B, uint32_t* nodes9);  // BVHGPUData, 9 words per node

// Parent of every reference node (root: 0xFFFFFFFF), from the skip links.
std::vector<uint32_t> reference_parents(const FlatBVH& B);

// Reachability links: for every reference node x, the nearest ancestor whose
// slab test is not implied by x's (0xFFFFFFFF: none).  A parent whose box
// contains its child's box (exact float compares of the stored corners)
// passes the reference's test whenever the child does, so the root-to-leaf
// test of a leaf needs only the leaf and the ancestors these links reach.
std::vector<uint32_t> reference_chain_links(const FlatBVH& B);

// Culling tree over the reference leaves: a binned-SAH binary tree whose
// leaves are the reference leaves (same triangles), each box the union of
// the reference leaf box and its triangles, padded by `pad`.  Same 8-word node
// format as FlatBVH; ref_leaf[i] is the reference node of culling leaf i.
// With split_cost > 0 a reference leaf of 3..32 triangles may instead enter
// as several primitives, groups of at most 2 of its triangles with their own
// padded boxes (kept when the surface-area estimate of its triangle tests
// drops by more than split_cost tests); such a wide leaf carries the group's
// triangle mask.
struct CullBVH {
    std::vector<uint32_t> nodes;
    std::vector<uint32_t> ref_leaf;
    uint32_t n_nodes = 0, depth = 0;
    // The same tree collapsed to 4 children per node, 32 words per node:
    // lo.x[4], lo.y[4], lo.z[4], hi.x[4], hi.y[4], hi.z[4], child[4], pad[4].
    // child: node index, kLeafFlag | leaf index, or kNoChild.  Leaves: 4 words
    // {first soup triangle, count, reference leaf node, triangle mask (bit j =
    // triangle first + j) or 0 for all of the leaf (or, with group nodes, the
    // leaf's group node + 1)}.
    std::vector<uint32_t> wide, wide_leaves;
    uint32_t n_wide = 0;
    uint32_t n_groups = 0;      // wide leaves that are triangle groups
    // Group nodes inside whole leaves (add_leaf_groups): the wide node format
    // with a triangle mask in place of each child; a whole wide leaf's word 3
    // is then its group node + 1 (0: none).
    std::vector<uint32_t> sub;
    uint32_t n_sub = 0;
    double sah = 0;             // surface-area cost of the 4-wide tree, in triangle tests
    double sah_split_pred = 0;  // the same predicted with every leaf of 3..32 triangles split
};
constexpr uint32_t kNoChild = 0xFFFFFFFFu;
// binary: also emit the binary (skip-link) culling tree.
void build_cull_bvh(const FlatBVH& ref, const std::vector<float>& soup_pos4, float pad, CullBVH& out,
                    float split_cost = 0.0f, bool binary = true);
constexpr double kBoxTestCost = 0.6;  // one box test, in triangle tests (measured on the AO kernel)
void add_leaf_groups(CullBVH& t, const std::vector<float>& soup_pos4, float pad);

}  // namespace fornos
