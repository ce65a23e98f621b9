/*
 * fornos_host.h -- host-side input producers for the baking hot path.
 *
 * These build what the reference's preprocessing layer hands to the GPU:
 * meshes (Mesh, Src/mesh.h:36-77) with their normal and tangent frames, the
 * texel set (MapUV -> CompressedMapUV, Src/compute.cpp:75-193,372-416) and
 * the hemisphere sample table (Src/math.h:339-382).  The meshes come from
 * seeded procedural generators instead of OBJ/PLY files.  All run on the
 * CPU and need no GPU.
 */
#ifndef FORNOS_HOST_H
#define FORNOS_HOST_H

#include "fornos_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fornos_mesh fornos_mesh;

/* Procedural generators.  Every generator emits the reference's Mesh model:
 * shared positions/texcoords/normals and one vertex per triangle corner, as
 * the OBJ loader builds it (Src/mesh.cpp:420-452). */
#define FORNOS_GEN_UV_SPHERE 0        /* low: radius a, b segments x c rings, analytic
                                         per-vertex normals, equirectangular UVs */
#define FORNOS_GEN_ICOSPHERE_WAVES 1  /* high: icosphere radius a, b subdivisions, radius
                                         displaced by 8 seeded sinusoids of amplitude c,
                                         frequency [2,16]; vertex normals recomputed */
#define FORNOS_GEN_HEIGHTFIELD 2      /* high: b x b quads over [0,1]^2, 6-octave Perlin fBm
                                         (amplitude a, lacunarity 2, gain 0.5) plus c seeded
                                         hemispherical bumps of radius [0.005,0.03]; vertex
                                         normals; planar UVs */
#define FORNOS_GEN_HEIGHTFIELD_LOW 3  /* low: b x b quads over [0,1]^2 displaced by the first
                                         2 octaves of the same fBm (same seed); vertex normals;
                                         planar UVs */
#define FORNOS_GEN_TORUS_KNOT 4       /* closed (p=3,q=5) torus-knot tube, b x c tube grid
                                         (along x around), displaced by 3-D fBm of amplitude a
                                         (a = 0: smooth); vertex normals; cylindrical UVs */

typedef struct fornos_gen_params {
    float a;
    uint32_t b;
    uint32_t c;
} fornos_gen_params;

fornos_status fornos_mesh_generate(int generator, uint64_t seed, const fornos_gen_params *p,
                                   fornos_mesh **out);
/* Copies a caller-described mesh into a library mesh. */
fornos_status fornos_mesh_from_view(const fornos_mesh_view *view, fornos_mesh **out);
/* Mesh::createCopy (Src/mesh.cpp:619-648). */
fornos_status fornos_mesh_copy(const fornos_mesh *m, fornos_mesh **out);
/* Pointers into the mesh's arrays, valid until the mesh changes or dies. */
fornos_status fornos_mesh_view_get(const fornos_mesh *m, fornos_mesh_view *out);
/* Mesh::computeVertexNormals (Src/mesh.cpp:650-678). */
fornos_status fornos_mesh_compute_vertex_normals(fornos_mesh *m);
/* Mesh::computeVertexNormalsAggressive (Src/mesh.cpp:680-721). */
fornos_status fornos_mesh_compute_vertex_normals_aggressive(fornos_mesh *m);
/* Mesh::computeFaceNormals (Src/mesh.cpp:627-648). */
fornos_status fornos_mesh_compute_face_normals(fornos_mesh *m);
/* Mesh::loadFile (Src/mesh.cpp:612-617): Wavefront .obj (loadWavefrontObj,
 * :235-456) or .ply (loadPly, :515-604) by extension; a parse error, an
 * index out of range or another extension is an error. */
fornos_status fornos_mesh_load(const char *path, fornos_mesh **out);
/* Mesh::computeTangentSpace (Src/mesh.cpp:724-768). */
fornos_status fornos_mesh_compute_tangent_space(fornos_mesh *m);
void fornos_mesh_destroy(fornos_mesh *m);

/* Texel set.  MapUV::fromMeshes(mesh, meshDirs, W, H) (Src/compute.cpp:358)
 * when dirs != NULL (the Smooth mapping method), MapUV::fromMesh otherwise,
 * followed by CompressedMapUV (Src/compute.cpp:372-416).  Returns
 * FORNOS_INVALID_ARG with "missing texture coordinates or normals" when a
 * triangle lacks them (Src/fornos.cpp:150-153). */
typedef struct fornos_texels fornos_texels;
fornos_status fornos_texels_create(const fornos_mesh *lo, const fornos_mesh *dirs, uint32_t width,
                                   uint32_t height, fornos_texels **out);
/* MapUV::fromMeshes_Hybrid (Src/compute.cpp:213-345, 365-370; the Hybrid
 * mapping method): the direction leans from the smooth direction toward the
 * surface normal within `edge` (FornosParameters_Shared::mappingEdge,
 * default 0.05) of a triangle edge.  edge = 0 gives fornos_texels_create. */
fornos_status fornos_texels_create_hybrid(const fornos_mesh *lo, const fornos_mesh *dirs, uint32_t width,
                                          uint32_t height, float edge, fornos_texels **out);
fornos_status fornos_texels_view_get(const fornos_texels *t, fornos_texel_view *out);
/* The uncompressed MapUV arrays (W*H*3 floats each); any pointer may be NULL. */
fornos_status fornos_texels_export_mapuv(const fornos_texels *t, float *positions,
                                         float *directions, float *normals, float *tangents,
                                         float *bitangents);
void fornos_texels_destroy(fornos_texels *t);

/* Host-only build of the reference-layout BVH of a mesh (the tree
 * fornos_scene_create uploads), exported as BVHGPUData nodes (9 words) and
 * the mesh triangle id of each soup triangle.  Call with nodes9 == NULL to
 * get the node count in *n_nodes; then nodes9 needs 9 * *n_nodes words and
 * soup_triangle_ids n_triangles words. */
fornos_status fornos_bvh_reference_export(const fornos_mesh *m, uint32_t max_tris_per_leaf,
                                          uint32_t max_depth, uint32_t *n_nodes,
                                          uint32_t *nodes9, uint32_t *soup_triangle_ids);

/* The same tree's parent links and the reachability links the hemisphere
 * kernels walk (each node's nearest ancestor whose slab test its own does
 * not imply, as the parent's stored box does not contain the node's);
 * 0xFFFFFFFF above the root.  n_nodes must be the tree's node count
 * (fornos_bvh_reference_export); either output may be NULL.  No reference
 * counterpart: the shaders test every node from the root down
 * (Shaders/ao_step1.comp:111-138). */
fornos_status fornos_bvh_reference_links(const fornos_mesh *m, uint32_t max_tris_per_leaf,
                                         uint32_t max_depth, uint32_t n_nodes, uint32_t *parent,
                                         uint32_t *links);

/* Hemisphere sample table: n_perm * S cosine-weighted directions (3 floats)
 * from hammersley(i, S, randX, randY) (Src/math.h:339-382).  (randX, randY)
 * pairs are drawn from a seeded std::mt19937 (two raw outputs per
 * permutation) instead of std::random_device; rand_out (2*n_perm, may be
 * NULL) receives them. */
fornos_status fornos_sample_table(uint32_t sample_count, uint32_t n_perm, uint32_t seed,
                                  float *out, uint32_t *rand_out);

#ifdef __cplusplus
}
#endif
#endif
