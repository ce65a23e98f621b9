/*
 * fornos_runner.h -- C ABI of the bake runner: the reference's
 * FornosParameters (Src/fornos.h:35-123), FornosRunner::start / pending /
 * run / currentTask (Src/fornos.h:139-149, Src/fornos.cpp:53-251) and the
 * FornosTask interface (Src/fornos.h:129-137: runStep, finish, progress,
 * name) over the device mapping, the solvers and the image export.
 *
 * Pacing: the reference's run() issues one runStep of the current task per
 * UI frame.  Here the first run() of a task submits all of its work to the
 * device without waiting (in a few parts with completion marks); later
 * run() calls poll the marks and return at once while the device works;
 * the run() that finds the task finished exports its map (the task's
 * finish) and moves to the next task.  Tasks run in the reference's order:
 * mesh mapping, height, position, normals, ambient occlusion, bent
 * normals, thickness.
 */
#ifndef FORNOS_RUNNER_H
#define FORNOS_RUNNER_H

#include "fornos_hip.h"
#include "fornos_host.h"

#ifdef __cplusplus
extern "C" {
#endif

/* NormalImport (Src/fornos.h:35) */
#define FORNOS_NORMALS_IMPORT 0
#define FORNOS_NORMALS_PER_FACE 1
#define FORNOS_NORMALS_PER_VERTEX 2
/* MeshMappingMethod (Src/fornos.h:36) */
#define FORNOS_MAPPING_SMOOTH 0
#define FORNOS_MAPPING_LOWPOLY_NORMALS 1
#define FORNOS_MAPPING_HYBRID 2

/* FornosParameters_Shared (Src/fornos.h:38-51).  Paths are UTF-8; an
 * empty or NULL hi-poly path bakes the low-poly mesh onto itself. */
typedef struct fornos_params_shared {
    const char *lo_poly_mesh_path;
    const char *hi_poly_mesh_path;
    int lo_poly_mesh_normal;   /* FORNOS_NORMALS_* */
    int hi_poly_mesh_normal;
    int bvh_tris_per_node;     /* 8 */
    int tex_width;             /* 2048 */
    int tex_height;            /* 2048 */
    int tex_dilation;          /* 16 */
    int ignore_backfaces;      /* 1 */
    int mapping;               /* FORNOS_MAPPING_* */
    float mapping_edge;        /* 0.05 */
    uint32_t sample_seed;      /* seed of the hemisphere sample tables (the
                                  reference draws them from std::random_device) */
} fornos_params_shared;

/* One map: the union of FornosParameters_Solver* (Src/fornos.h:53-113).
 * An empty or NULL output path computes the map without exporting it. */
typedef struct fornos_params_map {
    int enabled;
    const char *output_path;
    int sample_count;          /* AO / BN / thickness: 256 */
    float min_distance;        /* AO / BN / thickness: 0.01 */
    float max_distance;        /* AO / BN / thickness: 10; height: 0 */
    int tangent_space;         /* normals / BN: 1 */
    int normalize_output;      /* height: 1 */
} fornos_params_map;

/* FornosParameters (Src/fornos.h:114-123) */
typedef struct fornos_parameters {
    fornos_params_shared shared;
    fornos_params_map height;
    fornos_params_map positions;
    fornos_params_map normals;
    fornos_params_map ao;
    fornos_params_map bent_normals;
    fornos_params_map thickness;
} fornos_parameters;

/* The reference's member defaults (Src/fornos.h:38-113), every map disabled. */
void fornos_parameters_defaults(fornos_parameters *p);

typedef struct fornos_runner fornos_runner;
fornos_status fornos_runner_create(fornos_runner **out);
/* FornosRunner::start: loads and prepares the meshes, builds the texel set
 * on the device, the hi-poly BVH, the mesh mapping and one solver per
 * enabled map.  On failure returns non-zero and fornos_last_error() holds
 * the message (the reference's `errors` string, e.g. "Missing low poly
 * mesh").  Replaces any tasks still pending. */
fornos_status fornos_runner_start(fornos_runner *r, const fornos_parameters *p);
/* FornosRunner::pending */
int fornos_runner_pending(const fornos_runner *r);
/* FornosRunner::run: advances the current task without blocking on the
 * device (submits it, or polls it, or finishes it: export and removal). */
fornos_status fornos_runner_run(fornos_runner *r);
/* FornosRunner::run until no task is pending, waiting on the device between
 * polls (headless use). */
fornos_status fornos_runner_run_all(fornos_runner *r);
/* currentTask()->name() / progress(); *name = NULL when none is pending.
 * The name stays valid until the next run() or start(). */
fornos_status fornos_runner_current_task(fornos_runner *r, const char **name, float *progress);
/* Value range (min, max) of the float map a finished task exported (AO,
 * thickness, height); 0 when the runner has not exported that task.
 * task: FORNOS_SOLVER_* kind. */
int fornos_runner_exported_range(const fornos_runner *r, int task, float *minmax);
void fornos_runner_destroy(fornos_runner *r);

#ifdef __cplusplus
}
#endif
#endif
