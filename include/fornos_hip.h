/*
 * fornos_hip.h -- C ABI of the MI355X baking hot path.
 *
 * Plain C: POD structs, pointers and sizes, opaque handles, an int status on
 * every call.  No STL, GL or torch type crosses it.  Each entry point
 * replaces one method of the reference's GL-backed classes; the reference
 * interface it stands in for is cited (file:line, relative to the reference
 * root).  Host-side input producers (meshes, texel sets, sample tables) are
 * declared in fornos_host.h.
 *
 * Threading: a handle's work is queued on one HIP stream owned by its scene;
 * calls on one handle (and on handles sharing a scene) must be serialised by
 * the caller.  Distinct scenes may be driven from distinct threads.
 *
 * Ownership: as the reference's shared_ptr members (a solver holds its
 * MeshMapping, Src/solver_ao.h:93), a mapping holds a reference to its
 * scene, a solver to its mapping and a batch to its solvers.  *_destroy
 * drops the caller's reference; an object is freed when its last reference
 * goes, so handles may be destroyed in any order.
 */
#ifndef FORNOS_HIP_H
#define FORNOS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------ */
typedef int fornos_status;
#define FORNOS_OK 0
#define FORNOS_INVALID_ARG 1
#define FORNOS_OOM 2
#define FORNOS_HIP_ERROR 3
#define FORNOS_UNSUPPORTED 4

/* Thread-local message describing the last failed call on this thread.
 * Replaces the reference's logError strings (Src/logging.cpp:43-74). */
const char *fornos_last_error(void);

/* Library version string. */
const char *fornos_version(void);

/* Number of visible HIP devices (0 when none); never fails. */
int fornos_device_count(void);

/* ---- mesh view (the hi-poly Mesh of Src/mesh.h:36-77) ------------------- */
/* vertices: 3 uint32 per vertex {positionIndex, texcoordIndex, normalIndex}
 * (Mesh::Vertex, Src/mesh.h:38-43; UINT32_MAX = absent); triangles: 3 vertex
 * indices each (Mesh::Triangle, Src/mesh.h:45-50). */
typedef struct fornos_mesh_view {
    const float *positions;  /* 3 * n_positions */
    uint32_t n_positions;
    const float *normals;    /* 3 * n_normals */
    uint32_t n_normals;
    const float *texcoords;  /* 2 * n_texcoords (may be NULL) */
    uint32_t n_texcoords;
    const float *tangents;   /* 3 * n_vertices (may be NULL) */
    const float *bitangents; /* 3 * n_vertices (may be NULL) */
    const uint32_t *vertices;
    uint32_t n_vertices;
    const uint32_t *triangles;
    uint32_t n_triangles;
} fornos_mesh_view;

/* ---- scene: hi-poly mesh + BVH resident in HBM ------------------------- */
/* BVH layouts.  FORNOS_BVH_REFERENCE is the reference's binary tree
 * (Src/bvh.cpp:283-314, leaf <= max_tris_per_leaf or depth >= max_depth)
 * flattened in its DFS order with skip links (Src/meshmapping.cpp:62-125),
 * stored as 32-byte nodes. */
#define FORNOS_BVH_REFERENCE 0

typedef struct fornos_bvh_params {
    int layout;                 /* FORNOS_BVH_* */
    uint32_t max_tris_per_leaf; /* FornosParameters_Shared::bvhTrisPerNode, default 8 (Src/fornos.h:44) */
    uint32_t max_depth;         /* 8192 in the reference (Src/fornos.cpp:157) */
} fornos_bvh_params;

typedef struct fornos_scene_info {
    uint32_t n_nodes;
    uint32_t n_triangles;   /* soup triangles (= mesh triangles) */
    uint32_t n_leaves;
    uint32_t max_leaf_tris;
    uint32_t depth;
    double build_seconds;   /* host BVH build + flatten */
    size_t device_bytes;    /* HBM held by the scene */
    uint32_t ao_groups;     /* leaves of the AO tree that are triangle groups
                               (0: AO walks the culling tree of the others) */
    double sah;             /* surface-area cost of the 4-wide culling tree,
                               in triangle tests per ray */
    double sah_ao;          /* the same for the AO tree */
    uint32_t leaf_group_nodes;  /* whole leaves of the other passes' tree with
                                   group boxes inside (built with the AO tree) */
} fornos_scene_info;

typedef struct fornos_scene fornos_scene;

/* BVH::createBinary + fillMeshData + the GL uploads of MeshMapping::init
 * (Src/bvh.cpp:283, Src/meshmapping.cpp:128-162). */
fornos_status fornos_scene_create(const fornos_mesh_view *hi, const fornos_bvh_params *params,
                                  fornos_scene **out);
fornos_status fornos_scene_info_get(const fornos_scene *scene, fornos_scene_info *out);
/* Exports the node array in the reference's BVHGPUData layout
 * (Src/meshmapping.h:33-40: 6 floats min/max, start, end, jump; 9 words per
 * node; start/end are soup vertex indices), and the mesh triangle id of each
 * soup triangle (the flatten permutation).  Either pointer may be NULL. */
fornos_status fornos_scene_export_reference_nodes(const fornos_scene *scene, uint32_t *nodes9,
                                                  uint32_t *soup_triangle_ids);
/* Soup positions / normals as the reference's Vector4 arrays (4 floats per
 * soup vertex, Src/meshmapping.cpp:99-116).  Either pointer may be NULL. */
fornos_status fornos_scene_export_soup(const fornos_scene *scene, float *positions4,
                                       float *normals4);
void fornos_scene_destroy(fornos_scene *scene);

/* ---- mesh mapping (MeshMapping, Src/meshmapping.h:42-75) --------------- */
/* texels: CompressedMapUV arrays (Src/compute.h:245-259), 3 floats per texel.
 * normals/tangents/bitangents may be NULL when no tangent-space pass runs. */
typedef struct fornos_texel_view {
    uint32_t count;
    uint32_t width, height;
    const float *positions;
    const float *directions;
    const float *normals;
    const float *tangents;
    const float *bitangents;
    const uint32_t *indices;    /* MapUV index of each texel (may be NULL) */
} fornos_texel_view;

typedef struct fornos_meshmap fornos_meshmap;

/* MeshMapping::init(map, mesh, bvh, cullBackfaces) (Src/meshmapping.cpp:128-184). */
fornos_status fornos_meshmap_create(fornos_scene *scene, const fornos_texel_view *texels,
                                    int cull_backfaces, fornos_meshmap **out);
/* MeshMapping::init with the texel set built on the device: MapUV::fromMeshes
 * (Src/compute.cpp:358, :75-193) + CompressedMapUV (:372-416) + the uploads
 * of MeshMapping::init (Src/meshmapping.cpp:133-147), in one call.  lo: the
 * low-poly mesh (texcoords and normals required; per-vertex tangents and
 * bitangents optional); dirs: the mapping copy whose normals give the ray
 * directions (the Smooth method), or NULL to use lo's normals (fromMesh).
 * The texel set equals the host one (fornos_texels_create) bit for bit. */
fornos_status fornos_meshmap_create_from_mesh(fornos_scene *scene, const fornos_mesh_view *lo,
                                              const fornos_mesh_view *dirs, uint32_t width,
                                              uint32_t height, int cull_backfaces,
                                              fornos_meshmap **out);
/* The same with the Hybrid mapping method, MapUV::fromMeshes_Hybrid
 * (Src/compute.h:240, Src/compute.cpp:213-349) when edge > 0: ray
 * directions blended from dirs' normals toward lo's interpolated normal by
 * t = min(triangleDistance / edge, 1), tangent frame read by normal index,
 * bounding box clamped at 0 only.  edge == 0 is the call above.  Equals
 * fornos_texels_create_hybrid bit for bit. */
fornos_status fornos_meshmap_create_from_mesh_hybrid(fornos_scene *scene, const fornos_mesh_view *lo,
                                                     const fornos_mesh_view *dirs, uint32_t width,
                                                     uint32_t height, float edge, int cull_backfaces,
                                                     fornos_meshmap **out);
/* Readback of the texel set held by a mesh mapping (3 floats per texel each,
 * plus the MapUV index of each texel); any pointer may be NULL. */
fornos_status fornos_meshmap_read_texels(const fornos_meshmap *mm, float *positions,
                                         float *directions, float *normals, float *tangents,
                                         float *bitangents, uint32_t *indices);
/* MeshMapping::runStep (Src/meshmapping.cpp:186-218): one chunk of at most
 * 131072 texels; *done = 1 once every texel is mapped. */
fornos_status fornos_meshmap_run_step(fornos_meshmap *mm, int *done);
/* Every remaining texel in one launch.  *ms (may be NULL) receives the
 * kernel time measured with HIP events on the scene's stream. */
fornos_status fornos_meshmap_run_all(fornos_meshmap *mm, float *ms);
/* Rewinds progress to 0 so the next run_step/run_all maps every texel again
 * (the reference re-runs a pass only through a new init). */
fornos_status fornos_meshmap_reset(fornos_meshmap *mm);
/* Triangle test of the mapping.  FORNOS_PREDICATE_REFERENCE (default): the
 * reference's plane distance and fp64 barycentrics accepting b >= -1e-5
 * (Shaders/meshmapping.comp:67-86, meshmapping_nobackfaces.comp:60-100),
 * whose results the oracle checks bit for bit.  FORNOS_PREDICATE_WATERTIGHT:
 * the watertight test of Woop, Benthin and Wald (JCGT 2013): a ray through a
 * shared edge or vertex hits at least one of its triangles; same side
 * (culling) rule, same traversal order with a conservative box test (the
 * reference's rejects a box whose face lies in the plane that a ray with
 * a zero direction component travels in).  Hit ids differ from the
 * reference's only for rays passing near a triangle edge.  Takes effect at
 * the next run. */
#define FORNOS_PREDICATE_REFERENCE 0
#define FORNOS_PREDICATE_WATERTIGHT 1
fornos_status fornos_meshmap_set_predicate(fornos_meshmap *mm, int predicate);
/* Tree the mapping's passes walk with the reference predicate.
 * FORNOS_TRAVERSAL_WIDE (default): the 4-wide culling tree, nearest child
 * first, boxes skipped once entered beyond the best hit; a hit counts only
 * under the reference's reachability (its leaf's chain passes with the hit's
 * t; ties go to the smaller soup index, the reference's DFS order), and a
 * pass that cannot be decided so, or whose stack would overflow, is redone
 * on the reference tree.  FORNOS_TRAVERSAL_REFERENCE: the reference's own
 * stackless walk (Shaders/meshmapping_nobackfaces.comp:102-121).  Results
 * are identical; the watertight predicate always walks the reference tree.
 * No reference counterpart. */
#define FORNOS_TRAVERSAL_WIDE 0
#define FORNOS_TRAVERSAL_REFERENCE 1
/* stack: entries of the wide traversal's per-texel stack, 3..24, or 0 for
 * the default 24; a pass that would need more is redone on the reference
 * tree (results unchanged). */
fornos_status fornos_meshmap_set_traversal(fornos_meshmap *mm, int traversal, uint32_t stack);
/* Passes of the mapping runs since creation or the last reset that were
 * redone on the reference tree: *undecided for a hit whose reachability the
 * wide traversal could not decide, *overflow for a full stack.  Waits for
 * the scene's stream. */
fornos_status fornos_meshmap_redo_counts(fornos_meshmap *mm, uint32_t *undecided, uint32_t *overflow);
/* Asynchronous submission of the remaining mapping work as `parts` launches
 * (1..4096) with a completion mark after each; returns without waiting.
 * Replaces the reference's one-runStep-per-frame pacing
 * (Src/fornos.cpp:239-251, MeshMapping::runStep at Src/meshmapping.cpp:186-227):
 * same results, fewer launches. */
fornos_status fornos_meshmap_submit(fornos_meshmap *mm, uint32_t parts);
/* Non-blocking completion query: *progress = finished fraction of the
 * texels (by completed marks), *done = 1 once all submitted work has
 * finished on the device.  Without a submission: done when run_step/run_all
 * has covered all work and the stream is idle. */
fornos_status fornos_meshmap_poll(fornos_meshmap *mm, float *progress, int *done);
/* MeshMapping::progress (Src/meshmapping.h:48). */
fornos_status fornos_meshmap_progress(const fornos_meshmap *mm, float *out);
/* Readback of r_coords (t, b0, b1, b2 per texel) and r_tidx (soup vertex
 * index of the hit triangle, 0xFFFFFFFF on a miss), the contents of the
 * buffers MeshMapping::coords()/coords_tidx() expose (Src/meshmapping.h:50-51).
 * Either pointer may be NULL; count entries each. */
fornos_status fornos_meshmap_read(const fornos_meshmap *mm, float *coords4, uint32_t *tidx);
/* Number of texels (CompressedMapUV::positions.size()). */
fornos_status fornos_meshmap_texel_count(const fornos_meshmap *mm, uint32_t *out);
void fornos_meshmap_destroy(fornos_meshmap *mm);

/* ---- solvers (Src/solver_*.h) ------------------------------------------ */
#define FORNOS_SOLVER_HEIGHT 0        /* HeightSolver,      Src/solver_height.h      */
#define FORNOS_SOLVER_POSITION 1      /* PositionSolver,    Src/solver_position.h    */
#define FORNOS_SOLVER_NORMALS 2       /* NormalsSolver,     Src/solver_normals.h     */
#define FORNOS_SOLVER_AO 3            /* AmbientOcclusionSolver, Src/solver_ao.h     */
#define FORNOS_SOLVER_BENT_NORMALS 4  /* BentNormalsSolver, Src/solver_bentnormals.h */
#define FORNOS_SOLVER_THICKNESS 5     /* ThicknessSolver,   Src/solver_thickness.h   */

/* Union of the reference's solver Params structs. */
typedef struct fornos_solver_params {
    uint32_t sample_count;   /* AO / BN / thickness: sampleCount */
    float min_distance;      /* AO / BN / thickness: minDistance */
    float max_distance;      /* AO / BN / thickness / height: maxDistance */
    int tangent_space;       /* normals / BN: tangentSpace */
    int normalize_output;    /* height: normalizeOutput (used at export only) */
} fornos_solver_params;

typedef struct fornos_solver fornos_solver;

/* XSolver(params) + init(map, meshMapping) (e.g. Src/solver_ao.cpp:47-79).
 * sample_table: n_perm * sample_count directions, 3 floats each, indexed
 * (texel % n_perm) * sample_count + s (Shaders/ao_step1.comp:153); the
 * reference draws it with computeSamplesImportanceCosDir (Src/math.h:363-382,
 * n_perm = 4096); here the caller passes it as data.  Ignored (may be NULL)
 * for height/position/normals.  sample_count must be in [1, 4096].  The
 * library keeps host copies of the two most recent tables of at most 2^20
 * directions with their ray order (44 bytes a direction, until the process
 * ends): solvers created on an identical table reuse them. */
fornos_status fornos_solver_create(int kind, const fornos_solver_params *params,
                                   fornos_meshmap *mm, const float *sample_table,
                                   uint32_t n_perm, fornos_solver **out);
/* XSolver::runStep: one chunk of at most 131072 work items (rays for the
 * hemisphere solvers, texels otherwise), Src/solver_ao.cpp:81-131. */
fornos_status fornos_solver_run_step(fornos_solver *s, int *done);
/* All remaining work in one launch; *ms (may be NULL) = kernel time. */
fornos_status fornos_solver_run_all(fornos_solver *s, float *ms);
fornos_status fornos_solver_progress(const fornos_solver *s, float *out);
/* Rewinds progress to 0 so the next run_step/run_all repeats the whole pass
 * (same inputs, same results). */
fornos_status fornos_solver_reset(fornos_solver *s);
/* Pairs an AO solver with a bent-normals solver that casts the same rays
 * (same mesh mapping, sample table, sample count, min and max distance):
 * while both stand at the same point of their passes, the AO solver's
 * run_step / run_all / submit trace each ray once and produce both results
 * (hits in reachable leaves at t < maxDistance count for AO, any for bent
 * normals: Shaders/ao_step1.comp:157-165, bentnormals_step1.comp:170), and
 * the BN solver's pass is then already done.  Results are identical to the
 * two separate passes.  No reference counterpart (the reference runs the AO
 * and BN tasks one after the other, Src/fornos.cpp:174-198).  bn = NULL
 * unpairs.  Resetting the AO solver resets both. */
fornos_status fornos_solver_pair(fornos_solver *ao, fornos_solver *bn);
/* The same pair for a solver (runStep pacing of Src/solver_*.cpp). */
fornos_status fornos_solver_submit(fornos_solver *s, uint32_t parts);
fornos_status fornos_solver_poll(fornos_solver *s, float *progress, int *done);
/* XSolver::getResults (Src/solver_ao.cpp:133-138): copies count values into
 * caller memory: 1 float per texel (height, AO, thickness) or 3 (position,
 * normals, bent normals).  bytes must be at least that size. */
fornos_status fornos_solver_read(const fornos_solver *s, void *out, size_t bytes);
/* Per-texel ray counts of the last hemisphere run: occluded rays (AO, BN) or
 * rays that hit (thickness).  count uint32 entries. */
fornos_status fornos_solver_read_counts(const fornos_solver *s, uint32_t *out);
/* Kernel time of the last run_all of a hemisphere solver, split into the
 * ray launches (k_rays) and the per-texel reductions (k_reduce), from HIP
 * events on the scene's stream; *launches = number of k_rays launches. */
fornos_status fornos_solver_last_timing(const fornos_solver *s, float *rays_ms, float *reduce_ms,
                                        uint32_t *launches);
/* The pass of a hemisphere solver (AO, bent normals, thickness) for the
 * listed texels only: n distinct global texel indices into the solver's
 * texel set, in the order their rays are handed out (any order; the results
 * do not depend on it).  Their results, read with fornos_solver_read /
 * fornos_solver_read_counts, equal those of fornos_solver_run_all bit for bit
 * (each texel keeps its own sample permutation, texel % n_perm, as in
 * Shaders/ao_step1.comp:153); other texels' results are left as they are.
 * An AO solver paired with a bent-normals solver computes both.  Progress is
 * unchanged; *ms (may be NULL) = device time.  No reference counterpart (the
 * reference re-bakes a region only by baking the whole map,
 * Src/solver_ao.cpp:81-131); the denoised bake's per-tile fallback to the
 * plain bake uses it. */
fornos_status fornos_solver_run_texels(fornos_solver *s, const uint32_t *texels, uint32_t n, float *ms);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The pass of fornos_solver_run_all for a hemisphere
 * solver (AO, bent normals, thickness), with its results written into
 * caller-owned DEVICE memory on the solver's device -- out: count * (1 or 3)
 * floats, as fornos_solver_read returns them; counts: count uint32, as
 * fornos_solver_read_counts -- and its scratch (the ray counter and one
 * launch's bent-normal hit bits or thickness distances) in a fixed 64 MiB
 * device array of the library instead of memory allocated for the call.
 * Calls no hipMalloc or hipFree and waits for the pass to finish.  Launches
 * are cut to fit the array (thickness: 16383 texels per launch at
 * sample_count 1024), so a pass of any texel count runs; a solver of which one
 * texel alone needs more than the array is rejected with FORNOS_UNSUPPORTED.
 * Each device holds one such array for the life of the process and every call
 * on that device shares it: calls that overlap on one device (from several
 * threads, or on solvers of different scenes) race and give wrong results, so
 * the caller serialises them.  The solver's own results, progress and pairing
 * are left as they are (a paired bent-normals solver is not computed); rays
 * are handed out in the texel set's own order (the results are the same). */
fornos_status fornos_solver_run_all_noalloc(fornos_solver *s, float *out, uint32_t *counts);
/* Traversal work of the last run_all of a hemisphere solver, counted by an
 * instrumented run of the same kernel: node tests, triangle tests, rays. */
fornos_status fornos_solver_count_work(fornos_solver *s, uint64_t *node_tests,
                                       uint64_t *tri_tests, uint64_t *rays, uint64_t *stack_overflows);
/* All counters of the last fornos_solver_count_work run: box tests, triangle
 * tests, rays, stack overflows; then, summed over waves: walk steps, lanes
 * walking, leaf steps, lanes in leaf steps, refills, lanes refilled (their
 * ratios are the SIMD lane utilisation of each phase); then thickness rays
 * redone on the reference tree (a hit whose leaf the reference would prune
 * with the hit's own bound), and rays traced there from the start (a
 * direction component below 2^-40).  n <= 12. */
fornos_status fornos_solver_work_counters(const fornos_solver *s, uint64_t *counters, uint32_t n);
/* Per-texel flags of the last fornos_solver_count_work run, count uint32
 * entries: bit 0 = a ray of the texel overflowed the traversal stack and was
 * redone on the stackless binary culling tree, bit 1 = a thickness ray was
 * redone on the reference tree, bit 2 = a ray had a direction component below
 * 2^-40 and was traced on the reference tree.  These are the paths that take
 * rays off the 4-wide tree; tests point the oracle at the flagged texels.
 * No reference counterpart (the reference has one traversal). */
fornos_status fornos_solver_fallback_flags(const fornos_solver *s, uint32_t *flags);
/* Launch shape of the solver's ray kernel: persistent blocks of 256 lanes per
 * CU (1..16), idle lanes that trigger a wave's refill from its texel (1..64),
 * traversal stack entries per lane (16, 24 or 32; a ray that would overflow
 * it is redone on a stackless tree, so results do not depend on it); default
 * 8, 32, 16.  No reference counterpart (the shaders' dispatch shape is
 * fixed). */
fornos_status fornos_solver_set_launch(fornos_solver *s, uint32_t blocks_per_cu, uint32_t refill,
                                       uint32_t stack);
fornos_status fornos_solver_get_launch(const fornos_solver *s, uint32_t *blocks_per_cu, uint32_t *refill,
                                       uint32_t *stack);
/* Hand-out tile of a hemisphere solver's passes: texels a wave takes from
 * the pass's counter at once, 1..32, or 0 (default) for 256 / sampleCount
 * (1 at 256 samples and more; thickness at least 2).  A tile above 1 runs the ray kernel's variant
 * that keeps a wave's texel state in LDS.  Results do not depend on it.
 * get returns the tile in effect (the default resolved).  Other solvers:
 * FORNOS_INVALID_ARG.  No reference counterpart. */
fornos_status fornos_solver_set_tile(fornos_solver *s, uint32_t texels);
fornos_status fornos_solver_get_tile(const fornos_solver *s, uint32_t *texels);
/* Autotuner of a hemisphere solver's launch: times its pass over the first
 * `fraction` of its texels in the pass's order (at least 2^22 rays; the
 * paired AO + bent-normals traversal when paired) by coordinate descent --
 * stack depth {16, 24, 32}, then blocks per CU {4, 6, 8, 12}, then the
 * refill threshold {16, 24, 32, 40, 48}, then the hand-out tile {half, once,
 * twice the default} -- keeping on each axis the current value unless
 * another is at least 1 % faster; then times the ray order (texel order
 * against Morton order: whether the rays are sorted, once per mapping
 * result) on the same texels, every k-th of the set, and sets the other
 * order than the automatic one only when it is at least 1 % faster
 * (FORNOS_RAY_ORDER_TEXELS or _MORTON).  Results never depend on
 * any of these; the sampled texels are left with their pass results and the
 * solver's progress is unchanged.  No reference counterpart. */
fornos_status fornos_solver_autotune(fornos_solver *s, float fraction, uint32_t *blocks_per_cu, uint32_t *refill,
                                     uint32_t *stack, float *best_ms);
/* Order in which a hemisphere solver's passes hand out their texels' rays.
 * FORNOS_RAY_ORDER_MORTON: texels sorted by the Morton code of their hit
 * positions (built on the device once per mapping result), so the rays in
 * flight leave from a compact patch of the surface; on the 8.4M-triangle
 * knot (maxD 10) this cuts AO-512 and thickness-256 by 10 and 20 %.
 * FORNOS_RAY_ORDER_TEXELS: the texel set's own order (MapUV rows), as the
 * reference dispatches them (Shaders/ao_step1.comp:140-157).
 * FORNOS_RAY_ORDER_AUTO (default): Morton when maxDistance is at least a
 * quarter of the scene box's diagonal (rays that cross the scene), else the
 * texel order (short rays stay near a map row's texels, which already lie
 * together: heightfield2048's paired AO + BN pass is 3 % faster so).
 * Each texel's rays keep their sample permutation (texel % n_perm) and its
 * results their place, so results do not depend on the order; a run_step
 * covers the next 131072 rays of the order.  No reference counterpart. */
#define FORNOS_RAY_ORDER_TEXELS 0
#define FORNOS_RAY_ORDER_MORTON 1
#define FORNOS_RAY_ORDER_AUTO 2
fornos_status fornos_solver_set_ray_order(fornos_solver *s, int order);
fornos_status fornos_solver_get_ray_order(const fornos_solver *s, int *order);
/* The order the next pass will use: FORNOS_RAY_ORDER_TEXELS or _MORTON
 * (FORNOS_RAY_ORDER_AUTO resolved for this solver's maxDistance and scene). */
fornos_status fornos_solver_effective_ray_order(const fornos_solver *s, int *order);
void fornos_solver_destroy(fornos_solver *s);

/* ---- batched jobs -------------------------------------------------------- */
/* Several AO passes (each an AO solver with its own scene, texels and sample
 * count) run as one persistent k_rays launch over a device job table, plus
 * one reduction per job, captured into a HIP graph.  No reference
 * counterpart: the reference runs one solver step per UI frame
 * (Src/fornos.cpp:239-251); each job's results equal its solver's
 * fornos_solver_run_all bit for bit and are read with fornos_solver_read.
 * The batch holds a reference to each solver; their own streams must be
 * idle when fornos_batch_run is called. */
typedef struct fornos_batch fornos_batch;
fornos_status fornos_batch_create(fornos_solver *const *solvers, uint32_t n, fornos_batch **out);
/* Launch shape, as fornos_solver_set_launch. */
fornos_status fornos_batch_set_launch(fornos_batch *b, uint32_t blocks_per_cu, uint32_t refill, uint32_t stack);
fornos_status fornos_batch_get_launch(const fornos_batch *b, uint32_t *blocks_per_cu, uint32_t *refill,
                                      uint32_t *stack);
/* Autotuner: times launch shapes on the first `fraction` of each job's
 * texels (at least 2^22 rays of each job) by coordinate descent (stack depth {16,24,32}, then blocks per CU
 * {4,6,8,12}, then refill {16,24,32,40,48}, then the jobs' hand-out tiles: each job's solver tile
 * halved, kept or doubled), keeping on each axis the current value unless another is at least
 * 1 % faster.  Each job keeps its solver's ray order. */
fornos_status fornos_batch_autotune(fornos_batch *b, float fraction, uint32_t *blocks_per_cu,
                                    uint32_t *refill, uint32_t *stack, float *best_ms);
/* The jobs' hand-out tiles: each job's solver tile halved (-1), kept (0,
 * the default) or doubled (1), within 1..32; results do not depend on it. */
fornos_status fornos_batch_set_tile_shift(fornos_batch *b, int shift);
fornos_status fornos_batch_get_tile_shift(const fornos_batch *b, int *shift);
/* The hand-out tile job `s` (one of the batch's solvers) runs with. */
fornos_status fornos_batch_get_tile(const fornos_batch *b, const fornos_solver *s, uint32_t *tile);
/* Runs every job once (graph launch); *ms = wall time of the graph on the
 * device (HIP events). */
fornos_status fornos_batch_run(fornos_batch *b, float *ms);
void fornos_batch_destroy(fornos_batch *b);

#ifdef __cplusplus
}
#endif
#endif
