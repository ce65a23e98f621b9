/*
 * fornos_oracle.c -- CPU oracle for the Fornos baking hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is linked by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing that ships.  It is a plain-C restatement of the reference's
 * algorithm, written from the reference sources; every function cites the
 * reference file:line it follows (paths relative to the reference root).
 *
 * Arithmetic conventions (identical in the HIP product, which is what makes
 * bit-exact parity possible):
 *   - fp32 everywhere the reference uses float/vec3, fp64 where the mapping
 *     shaders use dvec3 (Shaders/meshmapping*.comp:48-62);
 *   - no contraction of a*b+c into FMA (built with -ffp-contract=off);
 *   - dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, left to right;
 *   - normalize(v) = v * (1.0f / sqrtf(dot(v,v))), with IEEE-correct sqrt and
 *     division.  This is Src/math.h:104-107 exactly; GLSL leaves normalize's
 *     rounding to the driver, so for the shaders this is a chosen convention;
 *   - GLSL min(x,y) = (y < x) ? y : x and max(x,y) = (x < y) ? y : x (the GLSL
 *     4.30 definitions), used in RayAABB; host code uses fminf/fmaxf as
 *     Src/math.h:110-124 does.
 *
 * Parity status of each part (what pins it) is listed in the oracle header
 * block of tests/test_oracle.py; functions without an independent pin are
 * marked "parity unpinned" below.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

typedef struct { float x, y, z; } v3;
typedef struct { float x, y; } v2;
typedef struct { double x, y, z; } d3;

static inline v3 mk3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul3s(v3 a, float k) { return mk3(a.x * k, a.y * k, a.z * k); }
static inline v3 div3s(v3 a, float k) { return mk3(a.x / k, a.y / k, a.z / k); }
static inline v3 neg3(v3 a) { return mk3(-a.x, -a.y, -a.z); }
static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* Src/math.h:101-107 */
static inline v3 norm3(v3 v) { return mul3s(v, 1.0f / sqrtf(dot3(v, v))); }
/* Src/math.h:110-124 */
static inline v3 min3h(v3 a, v3 b) { return mk3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static inline v3 max3h(v3 a, v3 b) { return mk3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
/* GLSL 4.30 min/max definitions */
static inline float gmin(float x, float y) { return (y < x) ? y : x; }
static inline float gmax(float x, float y) { return (x < y) ? y : x; }

static inline v3 ld3(const float *p, size_t i) { return mk3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void st3(float *p, size_t i, v3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
/* soup arrays use the reference's Vector4 (w = 0) stride: Src/meshmapping.cpp:64-68 */
static inline v3 ld4(const float *p, size_t i) { return mk3(p[4 * i], p[4 * i + 1], p[4 * i + 2]); }

/* ------------------------------------------------------------------------- */
/* Mesh normals and tangent frames                                            */
/* ------------------------------------------------------------------------- */

/* Src/mesh.cpp:650-678 Mesh::computeVertexNormals.  out_nrm has np entries;
 * each vertex's normal index becomes its position index. */
ORC_API void orc_vertex_normals(const float *pos, uint32_t np, const uint32_t *verts, uint32_t nv,
                                const uint32_t *tris, uint32_t ntri, float *out_nrm,
                                uint32_t *out_vert_nrm)
{
    memset(out_nrm, 0, sizeof(float) * 3 * np);
    for (uint32_t t = 0; t < ntri; ++t) {
        uint32_t p0i = verts[3 * tris[3 * t + 0]], p1i = verts[3 * tris[3 * t + 1]],
                 p2i = verts[3 * tris[3 * t + 2]];
        v3 p0 = ld3(pos, p0i), p1 = ld3(pos, p1i), p2 = ld3(pos, p2i);
        v3 n = norm3(cross3(sub3(p1, p0), sub3(p2, p0)));
        st3(out_nrm, p0i, add3(ld3(out_nrm, p0i), n));
        st3(out_nrm, p1i, add3(ld3(out_nrm, p1i), n));
        st3(out_nrm, p2i, add3(ld3(out_nrm, p2i), n));
    }
    for (uint32_t i = 0; i < np; ++i) st3(out_nrm, i, norm3(ld3(out_nrm, i)));
    for (uint32_t v = 0; v < nv; ++v) out_vert_nrm[v] = verts[3 * v];
}

/* Mesh::computeFaceNormals (Src/mesh.cpp:627-648): one normal per triangle,
 * every corner of the triangle re-pointed at it (last triangle wins). */
ORC_API void orc_face_normals(const float *pos, const uint32_t *verts, uint32_t nv, const uint32_t *tris,
                              uint32_t ntri, float *out_nrm, uint32_t *out_vert_nrm)
{
    for (uint32_t v = 0; v < nv; ++v) out_vert_nrm[v] = verts[3 * v + 2];
    for (uint32_t t = 0; t < ntri; ++t) {
        const uint32_t a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        v3 p0 = ld3(pos, verts[3 * a]), p1 = ld3(pos, verts[3 * b]), p2 = ld3(pos, verts[3 * c]);
        st3(out_nrm, t, norm3(cross3(sub3(p1, p0), sub3(p2, p0))));
        out_vert_nrm[a] = out_vert_nrm[b] = out_vert_nrm[c] = t;
    }
}

/* Exact-position key for the std::map<Vector3> of Src/mesh.cpp:682-697
 * (operator< at Src/mesh.cpp:38-41 makes -0 and +0 the same key). */
typedef struct { float x, y, z; uint32_t index; int used; } agg_slot;
static uint64_t agg_hash(float x, float y, float z)
{
    uint32_t a, b, c;
    if (x == 0.0f) x = 0.0f;
    if (y == 0.0f) y = 0.0f;
    if (z == 0.0f) z = 0.0f;
    memcpy(&a, &x, 4); memcpy(&b, &y, 4); memcpy(&c, &z, 4);
    uint64_t h = 1469598103934665603ull;
    h = (h ^ a) * 1099511628211ull; h = (h ^ b) * 1099511628211ull; h = (h ^ c) * 1099511628211ull;
    return h ^ (h >> 29);
}

/* Src/mesh.cpp:680-721 Mesh::computeVertexNormalsAggressive.  Corners with
 * bit-identical positions share one normal, the sum of their face normals in
 * triangle order.  Returns the number of normals written to out_nrm
 * (capacity 3*ntri). */
ORC_API uint32_t orc_vertex_normals_aggressive(const float *pos, const uint32_t *verts, uint32_t nv,
                                               const uint32_t *tris, uint32_t ntri, float *out_nrm,
                                               uint32_t *out_vert_nrm)
{
    (void)nv;
    uint32_t cap = 16;
    while (cap < 6 * ntri + 16) cap <<= 1;
    agg_slot *tab = (agg_slot *)calloc(cap, sizeof(agg_slot));
    uint32_t count = 0;
    for (uint32_t t = 0; t < ntri; ++t) {
        uint32_t vi[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
        v3 p[3];
        for (int k = 0; k < 3; ++k) p[k] = ld3(pos, verts[3 * vi[k]]);
        v3 n = norm3(cross3(sub3(p[1], p[0]), sub3(p[2], p[0])));
        for (int k = 0; k < 3; ++k) {
            uint64_t h = agg_hash(p[k].x, p[k].y, p[k].z) & (cap - 1);
            for (;;) {
                agg_slot *s = &tab[h];
                if (!s->used) {
                    s->used = 1; s->x = p[k].x; s->y = p[k].y; s->z = p[k].z; s->index = count;
                    st3(out_nrm, count, n);
                    out_vert_nrm[vi[k]] = count++;
                    break;
                }
                if (s->x == p[k].x && s->y == p[k].y && s->z == p[k].z) {
                    st3(out_nrm, s->index, add3(ld3(out_nrm, s->index), n));
                    out_vert_nrm[vi[k]] = s->index;
                    break;
                }
                h = (h + 1) & (cap - 1);
            }
        }
    }
    for (uint32_t i = 0; i < count; ++i) st3(out_nrm, i, norm3(ld3(out_nrm, i)));
    free(tab);
    return count;
}

/* Src/mesh.cpp:724-768 Mesh::computeTangentSpace: per-vertex accumulation of
 * the face tangent/bitangent directions, then Gram-Schmidt against the
 * vertex normal and b = cross(n, t). */
ORC_API void orc_tangent_space(const float *pos, const float *nrm, const float *uv,
                               const uint32_t *verts, uint32_t nv, const uint32_t *tris,
                               uint32_t ntri, float *out_t, float *out_b)
{
    memset(out_t, 0, sizeof(float) * 3 * nv);
    memset(out_b, 0, sizeof(float) * 3 * nv);
    for (uint32_t t = 0; t < ntri; ++t) {
        uint32_t a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        v3 p0 = ld3(pos, verts[3 * a]), p1 = ld3(pos, verts[3 * b]), p2 = ld3(pos, verts[3 * c]);
        v2 u0 = {uv[2 * verts[3 * a + 1]], uv[2 * verts[3 * a + 1] + 1]};
        v2 u1 = {uv[2 * verts[3 * b + 1]], uv[2 * verts[3 * b + 1] + 1]};
        v2 u2 = {uv[2 * verts[3 * c + 1]], uv[2 * verts[3 * c + 1] + 1]};
        v3 e0 = sub3(p1, p0), e1 = sub3(p2, p0);
        v2 t0 = {u1.x - u0.x, u1.y - u0.y}, t1 = {u2.x - u0.x, u2.y - u0.y};
        float r = 1.0f / (t0.x * t1.y - t1.x * t0.y);
        v3 sdir = mul3s(sub3(mul3s(e0, t1.y), mul3s(e1, t0.y)), r);
        v3 tdir = mul3s(sub3(mul3s(e1, t0.x), mul3s(e0, t1.x)), r);
        st3(out_t, a, add3(ld3(out_t, a), sdir)); st3(out_t, b, add3(ld3(out_t, b), sdir));
        st3(out_t, c, add3(ld3(out_t, c), sdir));
        st3(out_b, a, add3(ld3(out_b, a), tdir)); st3(out_b, b, add3(ld3(out_b, b), tdir));
        st3(out_b, c, add3(ld3(out_b, c), tdir));
    }
    for (uint32_t i = 0; i < nv; ++i) {
        v3 n = ld3(nrm, verts[3 * i + 2]);
        v3 t1 = ld3(out_t, i);
        v3 tt = norm3(sub3(t1, mul3s(n, dot3(n, t1))));
        st3(out_t, i, tt);
        st3(out_b, i, cross3(n, tt));
    }
}

/* ------------------------------------------------------------------------- */
/* Texel set: MapUV raster and CompressedMapUV                                */
/* ------------------------------------------------------------------------- */

/* Src/math.h:250-258 */
static inline v3 bary2(v2 p, v2 p0, v2 p1, v2 p2)
{
    v2 a = {p1.x - p0.x, p1.y - p0.y}, b = {p2.x - p0.x, p2.y - p0.y}, c = {p.x - p0.x, p.y - p0.y};
    float s = 1.0f / (a.x * b.y - b.x * a.y);
    float i = (c.x * b.y - b.x * c.y) * s;
    float j = (a.x * c.y - c.x * a.y) * s;
    return mk3(1.0f - i - j, i, j);
}

/* uint32_t(std::roundf(x)) as the reference's x86-64 build evaluates it: a
 * 64-bit truncating conversion whose low 32 bits are kept, so -1 becomes
 * 0xFFFFFFFF (Src/compute.cpp:136-139). */
static inline uint32_t round_u32(float x) { return (uint32_t)(int64_t)roundf(x); }

/* Src/compute.cpp:75-163 rasterTriangle + :165-193 createMapUV.
 * mesh: pos/nrm/uv/verts(pos,tex,nrm)/tris.  dnrm/dvert_nrm: normals of the
 * mapping copy (Smooth method) or NULL.  tan/bitan: per-vertex tangent frame
 * or NULL (zeros).  Outputs are W*H*3 floats each, zero-initialised here.
 * Returns 0 when a triangle lacks texcoords or normals (createMapUV returns
 * nullptr), 1 otherwise. */
ORC_API int orc_mapuv(const float *pos, const float *nrm, const float *uv, const uint32_t *verts,
                      const uint32_t *tris, uint32_t ntri, const float *dnrm,
                      const uint32_t *dvert_nrm, const float *tan, const float *bitan, uint32_t W,
                      uint32_t H, float *o_pos, float *o_dir, float *o_nrm, float *o_tan,
                      float *o_bit)
{
    size_t npx = (size_t)W * H;
    memset(o_pos, 0, sizeof(float) * 3 * npx); memset(o_dir, 0, sizeof(float) * 3 * npx);
    memset(o_nrm, 0, sizeof(float) * 3 * npx); memset(o_tan, 0, sizeof(float) * 3 * npx);
    memset(o_bit, 0, sizeof(float) * 3 * npx);
    v2 scale = {(float)W, (float)H};
    v2 pixsize = {1.0f / scale.x, 1.0f / scale.y};
    v2 halfpix = {pixsize.x * 0.5f, pixsize.y * 0.5f};
    for (uint32_t t = 0; t < ntri; ++t) {
        uint32_t vi[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
        for (int k = 0; k < 3; ++k)
            if (verts[3 * vi[k] + 1] == UINT32_MAX || verts[3 * vi[k] + 2] == UINT32_MAX) return 0;
        v3 p[3], n[3], d[3], tt[3], bb[3];
        v2 u[3];
        for (int k = 0; k < 3; ++k) {
            p[k] = ld3(pos, verts[3 * vi[k]]);
            u[k].x = uv[2 * verts[3 * vi[k] + 1]]; u[k].y = uv[2 * verts[3 * vi[k] + 1] + 1];
            n[k] = ld3(nrm, verts[3 * vi[k] + 2]);
            d[k] = dnrm ? ld3(dnrm, dvert_nrm[vi[k]]) : n[k];
            tt[k] = tan ? ld3(tan, vi[k]) : mk3(0, 0, 0);
            bb[k] = tan ? ld3(bitan, vi[k]) : mk3(0, 0, 0);
        }
        v2 s[3];
        for (int k = 0; k < 3; ++k) {
            s[k].x = (u[k].x - halfpix.x) * scale.x;
            s[k].y = (u[k].y - halfpix.y) * scale.y;
        }
        uint32_t xMin = round_u32(fminf(s[0].x, fminf(s[1].x, s[2].x)));
        uint32_t yMin = round_u32(fminf(s[0].y, fminf(s[1].y, s[2].y)));
        uint32_t xMax = round_u32(fmaxf(s[0].x, fmaxf(s[1].x, s[2].x)));
        uint32_t yMax = round_u32(fmaxf(s[0].y, fmaxf(s[1].y, s[2].y)));
        if (xMin > W) xMin = W;
        if (yMin > H) yMin = H;
        if (xMax > W) xMax = W;
        if (yMax > H) yMax = H;
        for (uint32_t y = yMin; y <= yMax; ++y) {
            for (uint32_t x = xMin; x <= xMax; ++x) {
                v2 q = {(float)x * pixsize.x + halfpix.x, (float)y * pixsize.y + halfpix.y};
                v3 b = bary2(q, u[0], u[1], u[2]);
                if (b.x >= -0.001f && b.x <= 1 && b.y >= -0.001f && b.y <= 1 && b.z >= -0.001f &&
                    b.z <= 1) {
                    size_t i = (size_t)y * W + x;
                    /* The reference writes past the end when y == H; the
                     * oracle drops such writes. */
                    if (i >= npx) continue;
                    st3(o_pos, i, add3(add3(mul3s(p[0], b.x), mul3s(p[1], b.y)), mul3s(p[2], b.z)));
                    st3(o_dir, i, norm3(add3(add3(mul3s(d[0], b.x), mul3s(d[1], b.y)), mul3s(d[2], b.z))));
                    st3(o_nrm, i, norm3(add3(add3(mul3s(n[0], b.x), mul3s(n[1], b.y)), mul3s(n[2], b.z))));
                    st3(o_tan, i, norm3(add3(add3(mul3s(tt[0], b.x), mul3s(tt[1], b.y)), mul3s(tt[2], b.z))));
                    st3(o_bit, i, norm3(add3(add3(mul3s(bb[0], b.x), mul3s(bb[1], b.y)), mul3s(bb[2], b.z))));
                }
            }
        }
    }
    return 1;
}

/* Src/compute.cpp:195-210: distance from p to the line through e0, e1, and
 * the smallest of the three edge-line distances. */
static float edge_distance(v3 e0, v3 e1, v3 p)
{
    v3 v = sub3(p, e0), e = sub3(e1, e0);
    float t = dot3(v, e) / dot3(e, e);
    v3 q = add3(e0, mul3s(e, t));
    v3 w = sub3(q, p);
    return sqrtf(dot3(w, w));
}

/* MapUV::fromMeshes_Hybrid (Src/compute.cpp:213-320, 365-370): the raster
 * of orc_mapuv with the Hybrid differences -- bounds (uint32)max(round(.), 0)
 * without the upper clamp, tangents indexed by normal index, and the ray
 * direction normalize(dsmooth * (1 - t) + n * t) with
 * t = min(triangleDistance / edge, 1). */
ORC_API int orc_mapuv_edge(const float *pos, const float *nrm, const float *uv, const uint32_t *verts,
                           const uint32_t *tris, uint32_t ntri, const float *dnrm,
                           const uint32_t *dvert_nrm, const float *tan, const float *bitan, uint32_t W,
                           uint32_t H, float edge, float *o_pos, float *o_dir, float *o_nrm, float *o_tan,
                           float *o_bit)
{
    size_t npx = (size_t)W * H;
    memset(o_pos, 0, sizeof(float) * 3 * npx); memset(o_dir, 0, sizeof(float) * 3 * npx);
    memset(o_nrm, 0, sizeof(float) * 3 * npx); memset(o_tan, 0, sizeof(float) * 3 * npx);
    memset(o_bit, 0, sizeof(float) * 3 * npx);
    v2 scale = {(float)W, (float)H};
    v2 pixsize = {1.0f / scale.x, 1.0f / scale.y};
    v2 halfpix = {pixsize.x * 0.5f, pixsize.y * 0.5f};
    for (uint32_t t = 0; t < ntri; ++t) {
        uint32_t vi[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
        for (int k = 0; k < 3; ++k)
            if (verts[3 * vi[k] + 1] == UINT32_MAX || verts[3 * vi[k] + 2] == UINT32_MAX) return 0;
        v3 p[3], n[3], d[3], tt[3], bb[3];
        v2 u[3];
        for (int k = 0; k < 3; ++k) {
            p[k] = ld3(pos, verts[3 * vi[k]]);
            u[k].x = uv[2 * verts[3 * vi[k] + 1]]; u[k].y = uv[2 * verts[3 * vi[k] + 1] + 1];
            n[k] = ld3(nrm, verts[3 * vi[k] + 2]);
            d[k] = dnrm ? ld3(dnrm, dvert_nrm[vi[k]]) : n[k];
            tt[k] = tan ? ld3(tan, verts[3 * vi[k] + 2]) : mk3(0, 0, 0);
            bb[k] = tan ? ld3(bitan, verts[3 * vi[k] + 2]) : mk3(0, 0, 0);
        }
        v2 s[3];
        for (int k = 0; k < 3; ++k) {
            s[k].x = (u[k].x - halfpix.x) * scale.x;
            s[k].y = (u[k].y - halfpix.y) * scale.y;
        }
        uint32_t xMin = (uint32_t)(int64_t)fmaxf(roundf(fminf(s[0].x, fminf(s[1].x, s[2].x))), 0.0f);
        uint32_t yMin = (uint32_t)(int64_t)fmaxf(roundf(fminf(s[0].y, fminf(s[1].y, s[2].y))), 0.0f);
        uint32_t xMax = (uint32_t)(int64_t)fmaxf(roundf(fmaxf(s[0].x, fmaxf(s[1].x, s[2].x))), 0.0f);
        uint32_t yMax = (uint32_t)(int64_t)fmaxf(roundf(fmaxf(s[0].y, fmaxf(s[1].y, s[2].y))), 0.0f);
        for (uint32_t y = yMin; y <= yMax; ++y) {
            for (uint32_t x = xMin; x <= xMax; ++x) {
                v2 q = {(float)x * pixsize.x + halfpix.x, (float)y * pixsize.y + halfpix.y};
                v3 b = bary2(q, u[0], u[1], u[2]);
                if (b.x >= -0.001f && b.x <= 1 && b.y >= -0.001f && b.y <= 1 && b.z >= -0.001f &&
                    b.z <= 1) {
                    size_t i = (size_t)y * W + x;
                    if (i >= npx) continue;  /* the reference writes past the end */
                    v3 pp = add3(add3(mul3s(p[0], b.x), mul3s(p[1], b.y)), mul3s(p[2], b.z));
                    v3 nn = norm3(add3(add3(mul3s(n[0], b.x), mul3s(n[1], b.y)), mul3s(n[2], b.z)));
                    float tdist = fminf(fminf(edge_distance(p[0], p[1], pp),
                                              fminf(edge_distance(p[0], p[2], pp), edge_distance(p[1], p[2], pp))) / edge,
                                        1.0f);
                    v3 ds = norm3(add3(add3(mul3s(d[0], b.x), mul3s(d[1], b.y)), mul3s(d[2], b.z)));
                    st3(o_pos, i, pp);
                    st3(o_dir, i, norm3(add3(mul3s(ds, 1.0f - tdist), mul3s(nn, tdist))));
                    st3(o_nrm, i, nn);
                    st3(o_tan, i, norm3(add3(add3(mul3s(tt[0], b.x), mul3s(tt[1], b.y)), mul3s(tt[2], b.z))));
                    st3(o_bit, i, norm3(add3(add3(mul3s(bb[0], b.x), mul3s(bb[1], b.y)), mul3s(bb[2], b.z))));
                }
            }
        }
    }
    return 1;
}

/* Src/compute.cpp:372-386: texels whose direction has dot(d,d) > 0.5, in
 * row-major order.  Writes their MapUV indices; returns the count. */
ORC_API uint32_t orc_compress(const float *dir, uint32_t npx, uint32_t *o_idx)
{
    uint32_t c = 0;
    for (uint32_t i = 0; i < npx; ++i) {
        v3 n = ld3(dir, i);
        if (dot3(n, n) > 0.5f) o_idx[c++] = i;
    }
    return c;
}

/* ------------------------------------------------------------------------- */
/* Hemisphere sample table                                                    */
/* ------------------------------------------------------------------------- */

/* Src/math.h:324-332 */
static inline uint32_t rev_bits(uint32_t bits)
{
    bits = (bits << 16u) | (bits >> 16u);
    bits = ((bits & 0x55555555u) << 1u) | ((bits & 0xAAAAAAAAu) >> 1u);
    bits = ((bits & 0x33333333u) << 2u) | ((bits & 0xCCCCCCCCu) >> 2u);
    bits = ((bits & 0x0F0F0F0Fu) << 4u) | ((bits & 0xF0F0F0F0u) >> 4u);
    bits = ((bits & 0x00FF00FFu) << 8u) | ((bits & 0xFF00FF00u) >> 8u);
    return bits;
}

/* Src/math.h:339-344 hammersley(i, N, randX, randY) and :363-382
 * computeSamplesImportanceCosDir(S, P, out), with the (randX, randY) pairs
 * given as data instead of drawn from std::random_device.  out is P*S*3. */
ORC_API void orc_samples(uint32_t S, uint32_t P, const uint32_t *rnd, float *out)
{
    for (uint32_t j = 0; j < P; ++j) {
        uint32_t rx = rnd[2 * j], ry = rnd[2 * j + 1];
        for (uint32_t i = 0; i < S; ++i) {
            float x = (float)i / (float)S + (float)(rx & 0xffffu) / (float)(1 << 16);
            float y = (float)(rev_bits(i) ^ ry) * 2.3283064365386963e-10f;
            float ux = x - (float)(long)x;
            float r = sqrtf(ux);
            float phi = (float)(2.0 * 3.14159265) * y;
            size_t o = (size_t)j * S + i;
            out[3 * o + 0] = r * cosf(phi);
            out[3 * o + 1] = r * sinf(phi);
            out[3 * o + 2] = sqrtf(1.0f - ux);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Reference binary BVH build (Src/bvh.cpp) and flatten (Src/meshmapping.cpp) */
/* ------------------------------------------------------------------------- */

typedef struct { v3 minv, maxv; } baabb;
static inline baabb baabb_empty(void) { baabb b; b.minv = mk3(FLT_MAX, FLT_MAX, FLT_MAX); b.maxv = mk3(-FLT_MAX, -FLT_MAX, -FLT_MAX); return b; }
static inline void baabb_add(baabb *b, v3 p) { b->minv = min3h(p, b->minv); b->maxv = max3h(p, b->maxv); }
/* Src/bvh.cpp:55-61 combine(a0, a1): min(a1.min, a0.min) */
static inline baabb baabb_comb(baabb a0, baabb a1) { baabb r; r.minv = min3h(a1.minv, a0.minv); r.maxv = max3h(a1.maxv, a0.maxv); return r; }
/* Src/bvh.cpp:63-68, including its surface-area term 2(xy + xz + xy). */
static inline float bcost(baabb a, int count)
{
    v3 s = sub3(a.maxv, a.minv);
    float area = 2.0f * (s.x * s.y + s.x * s.z + s.x * s.y);
    return (float)count * area;
}

typedef struct bnode {
    v3 center, size;            /* AABB as center / half size (Src/math.h:147-154) */
    size_t subtree;             /* subtreeTriangleCount */
    struct bnode *child;        /* NULL or array of 2 */
    uint32_t *tris; size_t ntris;
} bnode;

typedef struct { const float *pos; const uint32_t *verts; const uint32_t *tris; } bmesh;
static inline void tri_pts(const bmesh *m, uint32_t t, v3 *p0, v3 *p1, v3 *p2)
{
    *p0 = ld3(m->pos, m->verts[3 * m->tris[3 * t + 0]]);
    *p1 = ld3(m->pos, m->verts[3 * m->tris[3 * t + 1]]);
    *p2 = ld3(m->pos, m->verts[3 * m->tris[3 * t + 2]]);
}

/* Src/bvh.cpp:70-126 selectSplitFromBuckets (the SAH-like variant). */
static void select_split(const uint32_t bk[16], const baabb ba[16], size_t count, int *idx, float *cost)
{
    baabb right[15];
    right[14] = ba[15];
    for (int i = 13; i >= 0; --i) right[i] = baabb_comb(right[i + 1], ba[i + 1]);
    float best = FLT_MAX; int bi = 0;
    baabb left = baabb_empty();
    int cl = 0, cr = (int)count;
    for (int i = 0; i < 16; ++i) {
        cl += (int)bk[i]; cr -= (int)bk[i];
        left = baabb_comb(left, ba[i]);
        /* i == 15 reads aabbRightSide[15], one past the 15-entry array, in the
         * reference; its cost term is multiplied by countRight == 0. */
        float c = bcost(left, cl) + (i < 15 ? bcost(right[i], cr) : 0.0f);
        if (c < best) { best = c; bi = i; }
    }
    *idx = bi; *cost = best;
}

/* Src/bvh.cpp:128-203 findBestSplit */
static void find_split(const bmesh *m, const bnode *par, int *axis, float *split)
{
    baabb cb = baabb_empty();
    for (size_t k = 0; k < par->ntris; ++k) {
        v3 p0, p1, p2; tri_pts(m, par->tris[k], &p0, &p1, &p2);
        baabb_add(&cb, div3s(add3(add3(p0, p1), p2), 3.0f));
    }
    uint32_t bx[16] = {0}, by[16] = {0}, bz[16] = {0};
    baabb ax[16], ay[16], az[16];
    for (int i = 0; i < 16; ++i) ax[i] = ay[i] = az[i] = baabb_empty();
    v3 ext = sub3(cb.maxv, cb.minv);
    for (size_t k = 0; k < par->ntris; ++k) {
        v3 p0, p1, p2; tri_pts(m, par->tris[k], &p0, &p1, &p2);
        v3 c = div3s(add3(add3(p0, p1), p2), 3.0f);
        v3 d = sub3(c, cb.minv);
        v3 ijk = mk3(d.x / ext.x * 15.99f, d.y / ext.y * 15.99f, d.z / ext.z * 15.99f);
        size_t i = isnan(ijk.x) ? 0 : (size_t)ijk.x;
        size_t j = isnan(ijk.y) ? 0 : (size_t)ijk.y;
        size_t kk = isnan(ijk.z) ? 0 : (size_t)ijk.z;
        ++bx[i]; ++by[j]; ++bz[kk];
        baabb_add(&ax[i], p0); baabb_add(&ax[i], p1); baabb_add(&ax[i], p2);
        baabb_add(&ay[j], p0); baabb_add(&ay[j], p1); baabb_add(&ay[j], p2);
        baabb_add(&az[kk], p0); baabb_add(&az[kk], p1); baabb_add(&az[kk], p2);
    }
    int ix, iy, iz; float cx, cy, cz;
    select_split(bx, ax, par->ntris, &ix, &cx);
    select_split(by, ay, par->ntris, &iy, &cy);
    select_split(bz, az, par->ntris, &iz, &cz);
    if (cx <= cy && cx <= cz) {
        *axis = 0; *split = cb.minv.x + (cb.maxv.x - cb.minv.x) / 16.0f * (float)(ix + 1);
    } else if (cy <= cx && cy <= cz) {
        *axis = 1; *split = cb.minv.y + (cb.maxv.y - cb.minv.y) / 16.0f * (float)(iy + 1);
    } else {
        *axis = 2; *split = cb.minv.z + (cb.maxv.z - cb.minv.z) / 16.0f * (float)(iz + 1);
    }
}

/* Src/bvh.cpp:205-281 binaryDivisionBVH */
static void divide(const bmesh *m, size_t maxTris, size_t maxDepth, bnode *par, size_t depth)
{
    if (par->ntris <= maxTris || depth >= maxDepth) { par->subtree = par->ntris; return; }
    int axis; float split;
    find_split(m, par, &axis, &split);
    bnode *ch = (bnode *)calloc(2, sizeof(bnode));
    ch[0].tris = (uint32_t *)malloc(sizeof(uint32_t) * par->ntris);
    ch[1].tris = (uint32_t *)malloc(sizeof(uint32_t) * par->ntris);
    v3 minL = mk3(FLT_MAX, FLT_MAX, FLT_MAX), maxL = mk3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    v3 minR = minL, maxR = maxL;
    for (size_t k = 0; k < par->ntris; ++k) {
        uint32_t t = par->tris[k];
        v3 p0, p1, p2; tri_pts(m, t, &p0, &p1, &p2);
        v3 c = mul3s(add3(add3(p0, p1), p2), 1.0f / 3.0f);
        float cv = axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
        if (cv <= split) {
            ch[0].tris[ch[0].ntris++] = t;
            minL = min3h(minL, min3h(p0, min3h(p1, p2)));
            maxL = max3h(maxL, max3h(p0, max3h(p1, p2)));
        } else {
            ch[1].tris[ch[1].ntris++] = t;
            minR = min3h(minR, min3h(p0, min3h(p1, p2)));
            maxR = max3h(maxR, max3h(p0, max3h(p1, p2)));
        }
    }
    if (ch[0].ntris == 0 || ch[1].ntris == 0) {
        free(ch[0].tris); free(ch[1].tris); free(ch);
        par->subtree = par->ntris;
        return;
    }
    free(par->tris); par->tris = NULL; par->ntris = 0;
    par->child = ch;
    ch[0].center = mul3s(add3(minL, maxL), 0.5f); ch[0].size = mul3s(sub3(maxL, minL), 0.5f);
    ch[1].center = mul3s(add3(minR, maxR), 0.5f); ch[1].size = mul3s(sub3(maxR, minR), 0.5f);
    divide(m, maxTris, maxDepth, &ch[0], depth + 1);
    divide(m, maxTris, maxDepth, &ch[1], depth + 1);
    v3 aminL = sub3(ch[0].center, ch[0].size), amaxL = add3(ch[0].center, ch[0].size);
    v3 aminR = sub3(ch[1].center, ch[1].size), amaxR = add3(ch[1].center, ch[1].size);
    v3 amin = min3h(aminL, aminR), amax = max3h(amaxL, amaxR);
    par->center = mul3s(add3(amax, amin), 0.5f);
    par->size = mul3s(sub3(amax, amin), 0.5f);
    par->subtree = ch[0].subtree + ch[1].subtree;
}

static void free_tree(bnode *n)
{
    if (n->child) { free_tree(&n->child[0]); free_tree(&n->child[1]); free(n->child); }
    free(n->tris);
}

typedef struct {
    const bmesh *m; const float *nrm;
    uint32_t *nodes; uint32_t nn;        /* 9 words per node (BVHGPUData, Src/meshmapping.h:33-40) */
    float *spos, *snrm; uint32_t nsoup;  /* soup vertices, 4 floats each */
    uint32_t *stri;                      /* mesh triangle id per soup triangle */
} flat_ctx;

/* Src/meshmapping.cpp:62-125 fillMeshData */
static void flatten(flat_ctx *f, const bnode *b)
{
    if (!b->child && b->ntris == 0) return;
    if (b->child) {
        if (b->child[0].subtree > 0 && b->child[1].subtree == 0) { flatten(f, &b->child[0]); return; }
        if (b->child[1].subtree > 0 && b->child[0].subtree == 0) { flatten(f, &b->child[1]); return; }
    }
    uint32_t index = f->nn++;
    uint32_t *d = &f->nodes[9 * index];
    v3 mn = sub3(b->center, b->size), mx = add3(b->center, b->size);
    memcpy(&d[0], &mn.x, 4); memcpy(&d[1], &mn.y, 4); memcpy(&d[2], &mn.z, 4);
    memcpy(&d[3], &mx.x, 4); memcpy(&d[4], &mx.y, 4); memcpy(&d[5], &mx.z, 4);
    d[6] = f->nsoup;
    for (size_t k = 0; k < b->ntris; ++k) {
        uint32_t t = b->tris[k];
        f->stri[f->nsoup / 3] = t;
        for (int c = 0; c < 3; ++c) {
            uint32_t v = f->m->tris[3 * t + c];
            v3 p = ld3(f->m->pos, f->m->verts[3 * v]);
            v3 n = ld3(f->nrm, f->m->verts[3 * v + 2]);
            float *sp = &f->spos[4 * f->nsoup], *sn = &f->snrm[4 * f->nsoup];
            sp[0] = p.x; sp[1] = p.y; sp[2] = p.z; sp[3] = 0;
            sn[0] = n.x; sn[1] = n.y; sn[2] = n.z; sn[3] = 0;
            f->nsoup++;
        }
    }
    f->nodes[9 * index + 7] = f->nsoup;
    if (b->child) { flatten(f, &b->child[0]); flatten(f, &b->child[1]); }
    f->nodes[9 * index + 8] = f->nn;
}

/* Src/bvh.cpp:283-314 BVH::createBinary followed by fillMeshData.
 * Outputs: nodes (capacity 9*(2*ntri+1) words), soup positions and normals
 * (capacity 4*3*ntri floats each), soup_tri (capacity ntri).  Returns the node
 * count; *o_nsoup receives the soup vertex count. */
ORC_API uint32_t orc_bvh_build(const float *pos, uint32_t np, const float *nrm,
                               const uint32_t *verts, const uint32_t *tris, uint32_t ntri,
                               uint32_t max_tris, uint32_t max_depth, uint32_t *o_nodes,
                               float *o_spos, float *o_snrm, uint32_t *o_stri, uint32_t *o_nsoup)
{
    bmesh m = {pos, verts, tris};
    v3 mins = mk3(FLT_MAX, FLT_MAX, FLT_MAX), maxs = mk3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (uint32_t i = 0; i < np; ++i) { v3 p = ld3(pos, i); mins = min3h(mins, p); maxs = max3h(maxs, p); }
    bnode root; memset(&root, 0, sizeof(root));
    root.center = mul3s(add3(maxs, mins), 0.5f);
    root.size = mul3s(sub3(maxs, mins), 0.5f);
    root.ntris = ntri;
    root.tris = (uint32_t *)malloc(sizeof(uint32_t) * (ntri ? ntri : 1));
    for (uint32_t i = 0; i < ntri; ++i) root.tris[i] = i;
    divide(&m, max_tris, max_depth, &root, 0);
    flat_ctx f = {&m, nrm, o_nodes, 0, o_spos, o_snrm, 0, o_stri};
    flatten(&f, &root);
    free_tree(&root);
    *o_nsoup = f.nsoup;
    return f.nn;
}

/* ------------------------------------------------------------------------- */
/* Shaders                                                                    */
/* ------------------------------------------------------------------------- */

typedef struct { const uint32_t *w; } nodes_t;
static inline v3 node_min(const uint32_t *n, uint32_t i) { v3 r; memcpy(&r.x, &n[9 * i], 4); memcpy(&r.y, &n[9 * i + 1], 4); memcpy(&r.z, &n[9 * i + 2], 4); return r; }
static inline v3 node_max(const uint32_t *n, uint32_t i) { v3 r; memcpy(&r.x, &n[9 * i + 3], 4); memcpy(&r.y, &n[9 * i + 4], 4); memcpy(&r.z, &n[9 * i + 5], 4); return r; }

/* Shaders/ao_step1.comp:43-52 (same text in every traversal shader) */
static inline float ray_aabb(v3 o, v3 d, v3 mins, v3 maxs)
{
    v3 t1 = mk3((mins.x - o.x) / d.x, (mins.y - o.y) / d.y, (mins.z - o.z) / d.z);
    v3 t2 = mk3((maxs.x - o.x) / d.x, (maxs.y - o.y) / d.y, (maxs.z - o.z) / d.z);
    v3 tmin = mk3(gmin(t1.x, t2.x), gmin(t1.y, t2.y), gmin(t1.z, t2.z));
    v3 tmax = mk3(gmax(t1.x, t2.x), gmax(t1.y, t2.y), gmax(t1.z, t2.z));
    float a = gmax(tmin.x, gmax(tmin.y, tmin.z));
    float b = gmin(tmax.x, gmin(tmax.y, tmax.z));
    return (b >= 0 && a <= b) ? a : FLT_MAX;
}

/* Shaders/meshmapping.comp:50-64 (fp64 barycentrics, result narrowed to vec3) */
static inline v3 bary_f64(v3 pf, v3 af, v3 bf, v3 cf)
{
    d3 p = {pf.x, pf.y, pf.z}, a = {af.x, af.y, af.z}, b = {bf.x, bf.y, bf.z}, c = {cf.x, cf.y, cf.z};
    d3 v0 = {b.x - a.x, b.y - a.y, b.z - a.z}, v1 = {c.x - a.x, c.y - a.y, c.z - a.z},
       v2 = {p.x - a.x, p.y - a.y, p.z - a.z};
    double d00 = v0.x * v0.x + v0.y * v0.y + v0.z * v0.z;
    double d01 = v0.x * v1.x + v0.y * v1.y + v0.z * v1.z;
    double d11 = v1.x * v1.x + v1.y * v1.y + v1.z * v1.z;
    double d20 = v2.x * v0.x + v2.y * v0.y + v2.z * v0.z;
    double d21 = v2.x * v1.x + v2.y * v1.y + v2.z * v1.z;
    double denom = d00 * d11 - d01 * d01;
    double y = (d11 * d20 - d01 * d21) / denom;
    double z = (d00 * d21 - d01 * d20) / denom;
    return mk3((float)(1.0 - y - z), (float)y, (float)z);
}

#define BARY_MIN (-1e-5f)
#define BARY_MAX (1.0f)

/* mode 0: Shaders/meshmapping.comp:67-86 (|nd| > 0); 1: forward, nd > 0
 * (meshmapping_nobackfaces.comp:64-83); 2: backward, nd < 0 (:123-142).
 * maxdist bounds t for modes 1 and 2 only. */
static inline int map_tri(v3 o, v3 d, v3 a, v3 b, v3 c, int mode, float maxdist, float *t_out, v3 *bc)
{
    v3 n = norm3(cross3(sub3(b, a), sub3(c, a)));
    float nd = dot3(d, n);
    int ok = mode == 0 ? (fabsf(nd) > 0) : (mode == 1 ? nd > 0 : nd < 0);
    if (!ok) return 0;
    float pn = dot3(o, n);
    float t = (dot3(a, n) - pn) / nd;
    if (mode == 0 ? !(t >= 0) : !(t >= 0 && t < maxdist)) return 0;
    v3 p = add3(o, mul3s(d, t));
    v3 w = bary_f64(p, a, b, c);
    if (w.x >= BARY_MIN && w.y >= BARY_MIN && w.y <= BARY_MAX && w.z >= BARY_MIN && w.z <= BARY_MAX) {
        *t_out = t; *bc = w; return 1;
    }
    return 0;
}

/* Shaders/meshmapping.comp:107-137 raycastBVH (non-culling) */
static float map_bvh_nocull(v3 o, v3 d, float mint, const uint32_t *nodes, uint32_t nn,
                            const float *spos, uint32_t *o_idx, v3 *o_bc)
{
    uint32_t i = 0;
    while (i < nn) {
        float da = ray_aabb(o, d, node_min(nodes, i), node_max(nodes, i));
        if (da < mint) {
            float lmin = FLT_MAX; uint32_t ridx = 0; v3 rbc = mk3(0, 0, 0);
            for (uint32_t t = nodes[9 * i + 6]; t < nodes[9 * i + 7]; t += 3) {
                float tt; v3 bc;
                if (map_tri(o, d, ld4(spos, t), ld4(spos, t + 1), ld4(spos, t + 2), 0, 0, &tt, &bc) &&
                    tt >= 0 && tt < lmin) {
                    lmin = tt; ridx = t; rbc = bc;
                }
            }
            if (lmin < mint) { mint = lmin; *o_idx = ridx; *o_bc = rbc; }
            ++i;
        } else {
            i = nodes[9 * i + 8];
        }
    }
    return mint;
}

/* Shaders/meshmapping_nobackfaces.comp:102-121 and :161-180 */
static void map_bvh_cull(v3 o, v3 d, int mode, float *cur, const uint32_t *nodes, uint32_t nn,
                         const float *spos, uint32_t *o_idx, v3 *o_bc)
{
    uint32_t i = 0;
    while (i < nn) {
        float da = ray_aabb(o, d, node_min(nodes, i), node_max(nodes, i));
        if (da < *cur) {
            for (uint32_t t = nodes[9 * i + 6]; t < nodes[9 * i + 7]; t += 3) {
                float tt; v3 bc;
                if (map_tri(o, d, ld4(spos, t), ld4(spos, t + 1), ld4(spos, t + 2), mode, *cur, &tt, &bc)) {
                    *cur = tt; *o_idx = t; *o_bc = bc;
                }
            }
            ++i;
        } else {
            i = nodes[9 * i + 8];
        }
    }
}

/* Shaders/meshmapping.comp:139-162 and meshmapping_nobackfaces.comp:182-199.
 * pix_p/pix_d: 3 floats per texel (Pix_GPUData p, d).  Texels [begin, end). */
ORC_API void orc_meshmapping(const float *pix_p, const float *pix_d, uint32_t begin, uint32_t end,
                             const float *spos, const uint32_t *nodes, uint32_t nn, int cull,
                             float *o_coords, uint32_t *o_tidx, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t g = begin; g < (int64_t)end; ++g) {
        v3 p = ld3(pix_p, g), d = ld3(pix_d, g);
        uint32_t tidx = 4294967295u; v3 bc = mk3(0, 0, 0); float t = FLT_MAX;
        if (cull) {
            map_bvh_cull(p, d, 1, &t, nodes, nn, spos, &tidx, &bc);
            map_bvh_cull(p, neg3(d), 2, &t, nodes, nn, spos, &tidx, &bc);
        } else {
            float r = map_bvh_nocull(p, d, t, nodes, nn, spos, &tidx, &bc);
            t = gmin(t, r);
            r = map_bvh_nocull(p, neg3(d), t, nodes, nn, spos, &tidx, &bc);
            t = gmin(t, r);
        }
        o_coords[4 * g + 0] = t; o_coords[4 * g + 1] = bc.x; o_coords[4 * g + 2] = bc.y; o_coords[4 * g + 3] = bc.z;
        o_tidx[g] = tidx;
    }
}

/* Shaders/ao_step0.comp:39-56: ray origin, direction and tangent frame. */
static inline void ray_basis(const float *coords, const uint32_t *tidx, const float *spos,
                             const float *snrm, uint32_t g, v3 *o, v3 *d, v3 *tx, v3 *ty)
{
    const float *c = &coords[4 * (size_t)g];
    uint32_t t = tidx[g];
    v3 p0 = ld4(spos, t), p1 = ld4(spos, t + 1), p2 = ld4(spos, t + 2);
    *o = add3(add3(mul3s(p0, c[1]), mul3s(p1, c[2])), mul3s(p2, c[3]));
    v3 n0 = ld4(snrm, t), n1 = ld4(snrm, t + 1), n2 = ld4(snrm, t + 2);
    *d = norm3(add3(add3(mul3s(n0, c[1]), mul3s(n1, c[2])), mul3s(n2, c[3])));
    v3 dd = *d;
    *ty = norm3(fabsf(dd.x) > fabsf(dd.y) ? mk3(dd.z, 0, -dd.x) : mk3(0, dd.z, -dd.y));
    *tx = cross3(dd, *ty);
}

/* Shaders/ao_step1.comp:109-147 (fp32 barycentrics, no tolerance) */
static inline float hemi_tri(v3 o, v3 d, v3 a, v3 b, v3 c)
{
    v3 n = norm3(cross3(sub3(b, a), sub3(c, a)));
    float nd = dot3(d, n);
    if (fabsf(nd) > 0) {
        float pn = dot3(o, n);
        float t = (dot3(a, n) - pn) / nd;
        if (t >= 0) {
            v3 p = add3(o, mul3s(d, t));
            v3 v0 = sub3(b, a), v1 = sub3(c, a), v2 = sub3(p, a);
            float d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1);
            float d20 = dot3(v2, v0), d21 = dot3(v2, v1);
            float denom = d00 * d11 - d01 * d01;
            float y = (d11 * d20 - d01 * d21) / denom;
            float z = (d00 * d21 - d01 * d20) / denom;
            float x = 1.0f - y - z;
            if (x >= 0 && y >= 0 && y <= 1 && z >= 0 && z <= 1) return t;
        }
    }
    return FLT_MAX;
}

/* Shaders/ao_step1.comp:166-193 raycastBVH (identical in bentnormals_step1
 * and thick_step1).  Optional counters record node and triangle tests. */
static float hemi_bvh(v3 o, v3 d, float mind, float maxd, const uint32_t *nodes, uint32_t nn,
                      const float *spos, uint64_t *cnt_nodes, uint64_t *cnt_tris)
{
    float mint = FLT_MAX;
    uint32_t i = 0;
    uint64_t cn = 0, ct = 0;
    while (i < nn) {
        ++cn;
        float da = ray_aabb(o, d, node_min(nodes, i), node_max(nodes, i));
        if (da < mint && da < maxd) {
            float lmin = FLT_MAX;
            for (uint32_t t = nodes[9 * i + 6]; t < nodes[9 * i + 7]; t += 3) {
                ++ct;
                float tt = hemi_tri(o, d, ld4(spos, t), ld4(spos, t + 1), ld4(spos, t + 2));
                if (tt >= mind && tt < lmin) lmin = tt;
            }
            if (lmin < mint) mint = lmin;
            ++i;
        } else {
            i = nodes[9 * i + 8];
        }
    }
    if (cnt_nodes) *cnt_nodes += cn;
    if (cnt_tris) *cnt_tris += ct;
    return mint;
}

/* Hemisphere passes.  kind 0 = AO (ao_step0/1/2), 1 = bent normals
 * (ao_step0 + bentnormals_step1/2, before tangentspace), 2 = thickness
 * (ao_step0 + thick_step1/2).  Texels [begin, end) of the compressed set.
 * samples: P*S*3 floats.  out: 1 float per texel (AO, thickness) or 3 (BN),
 * at index (g - begin).  Texels whose mapping missed (tidx == 0xFFFFFFFF),
 * which the reference reads out of bounds for, get AO 1, BN (0,0,0) and
 * thickness maxD.  Optional out_count receives, per texel, the number of
 * occluded rays (AO, BN) or of rays that hit (thickness). */
ORC_API void orc_hemisphere(int kind, uint32_t S, uint32_t P, float mind, float maxd,
                            const float *samples, const float *coords, const uint32_t *tidx,
                            uint32_t begin, uint32_t end, const float *spos, const float *snrm,
                            const uint32_t *nodes, uint32_t nn, float *out, uint32_t *out_count,
                            uint64_t *stats, int nthreads)
{
    uint64_t tot_nodes = 0, tot_tris = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : tot_nodes, tot_tris)
#endif
    for (int64_t g = begin; g < (int64_t)end; ++g) {
        size_t oi = (size_t)(g - begin);
        uint32_t cnt = 0;
        if (tidx[g] == 4294967295u) {
            if (kind == 1) { out[3 * oi] = 0; out[3 * oi + 1] = 0; out[3 * oi + 2] = 0; }
            else out[oi] = kind == 0 ? 1.0f : maxd;
            if (out_count) out_count[oi] = 0;
            continue;
        }
        v3 o, d, tx, ty;
        ray_basis(coords, tidx, spos, snrm, (uint32_t)g, &o, &d, &tx, &ty);
        if (kind == 2) d = neg3(d);          /* thick_step1.comp:144 */
        float facc = 0; v3 vacc = mk3(0, 0, 0);
        uint64_t ln = 0, lt = 0;
        for (uint32_t s = 0; s < S; ++s) {
            size_t sidx = (size_t)((uint32_t)g % P) * S + s;    /* ao_step1.comp:153 */
            v3 rs = ld3(samples, sidx);
            v3 dir = norm3(add3(add3(mul3s(tx, rs.x), mul3s(ty, rs.y)), mul3s(d, rs.z)));
            float t = hemi_bvh(o, dir, mind, maxd, nodes, nn, spos, stats ? &ln : NULL, stats ? &lt : NULL);
            if (kind == 0) {                 /* ao_step1.comp:158-165 */
                float r = (t != FLT_MAX && t < maxd) ? 1.0f : 0.0f;
                facc += r; cnt += (uint32_t)r;
            } else if (kind == 1) {          /* bentnormals_step1.comp:170 */
#ifdef ORC_MUTANT_BN_MAXD
                /* planted mistake (mutant builds for the pin tests only) */
                if (t != FLT_MAX && t < maxd) ++cnt; else vacc = add3(vacc, dir);
#else
                if (t != FLT_MAX) ++cnt; else vacc = add3(vacc, dir);
#endif
            } else {                         /* thick_step1.comp:158 */
#ifdef ORC_MUTANT_THICK_CLAMP
                /* planted mistake (mutant builds for the pin tests only) */
                if (t != FLT_MAX && t > maxd) t = maxd;
#endif
                facc += (t != FLT_MAX) ? t : maxd;
                if (t != FLT_MAX) ++cnt;
            }
        }
        if (kind == 0) out[oi] = 1.0f - facc / (float)S;      /* ao_step2.comp:24-29 */
        else if (kind == 2) out[oi] = facc / (float)S;        /* thick_step2.comp:24-29 */
        else {                                                /* bentnormals_step2.comp:25-34 */
            v3 nn3 = norm3(vacc);
            out[3 * oi] = nn3.x; out[3 * oi + 1] = nn3.y; out[3 * oi + 2] = nn3.z;
        }
        if (out_count) out_count[oi] = cnt;
        tot_nodes += ln; tot_tris += lt;
    }
    if (stats) { stats[0] = tot_nodes; stats[1] = tot_tris; }
}

/* Shaders/tangentspace.comp:22-41, in place on 3-float values.  n, t, b are
 * the low-poly per-texel frame (PixelT). */
ORC_API void orc_tangentspace(const float *n_, const float *t_, const float *b_, uint32_t begin,
                              uint32_t end, float *vals)
{
    for (uint32_t g = begin; g < end; ++g) {
        v3 v = ld3(vals, g - begin), n = ld3(n_, g), t = ld3(t_, g), b = ld3(b_, g);
        v3 d0 = mk3(n.z * b.y - n.y * b.z, n.x * b.z - n.z * b.x, n.y * b.x - n.x * b.y);
        v3 d1 = mk3(t.z * n.y - t.y * n.z, t.x * n.z - n.x * t.z, n.x * t.y - t.x * n.y);
        v3 d2 = mk3(t.y * b.z - t.z * b.y, t.z * b.x - t.x * b.z, t.x * b.y - t.y * b.x);
        st3(vals, g - begin, norm3(mk3(dot3(v, d0), dot3(v, d1), dot3(v, d2))));
    }
}

/* Shaders/heights.comp:13-19 */
ORC_API void orc_height(const float *coords, uint32_t n, float *out)
{
    for (uint32_t g = 0; g < n; ++g) { float h = coords[4 * g]; out[g] = h != FLT_MAX ? h : 0; }
}

/* Shaders/positions.comp:17-32 (kind 0) and normals.comp:13-28 (kind 1).
 * Missed texels (tidx == 0xFFFFFFFF, read out of bounds by the reference)
 * get (0,0,0). */
ORC_API void orc_texel_attr(int kind, const float *coords, const uint32_t *tidx, uint32_t n,
                            const float *soup, float *out)
{
    for (uint32_t g = 0; g < n; ++g) {
        uint32_t t = tidx[g];
        if (t == 4294967295u) { st3(out, g, mk3(0, 0, 0)); continue; }
        const float *c = &coords[4 * (size_t)g];
        v3 a = ld4(soup, t), b = ld4(soup, t + 1), cc = ld4(soup, t + 2);
        v3 v = add3(add3(mul3s(a, c[1]), mul3s(b, c[2])), mul3s(cc, c[3]));
        st3(out, g, kind == 1 ? norm3(v) : v);
    }
}

/* Brute force closest hit over all soup triangles with the mapping predicate
 * of the culling variant; used to pin the BVH traversal (its result must not
 * depend on the BVH).  Ties keep the earlier soup triangle; the backward pass
 * must beat the forward one strictly. */
ORC_API void orc_map_bruteforce(const float *pix_p, const float *pix_d, uint32_t n,
                                const float *spos, uint32_t nsoup, int cull, float *o_coords,
                                uint32_t *o_tidx)
{
    for (uint32_t g = 0; g < n; ++g) {
        v3 p = ld3(pix_p, g), d = ld3(pix_d, g);
        float cur = FLT_MAX; uint32_t idx = 4294967295u; v3 bc = mk3(0, 0, 0);
        for (int pass = 0; pass < 2; ++pass) {
            v3 dd = pass ? neg3(d) : d;
            float pass_best = FLT_MAX; uint32_t pidx = 0; v3 pbc = bc;
            for (uint32_t t = 0; t < nsoup; t += 3) {
                float tt; v3 w;
                int mode = cull ? (pass ? 2 : 1) : 0;
                if (map_tri(p, dd, ld4(spos, t), ld4(spos, t + 1), ld4(spos, t + 2), mode,
                            cull ? cur : FLT_MAX, &tt, &w)) {
                    if (cull) { cur = tt; idx = t; bc = w; }
                    else if (tt < pass_best) { pass_best = tt; pidx = t; pbc = w; }
                }
            }
            if (!cull && pass_best < cur) { cur = pass_best; idx = pidx; bc = pbc; }
        }
        o_coords[4 * g] = cur; o_coords[4 * g + 1] = bc.x; o_coords[4 * g + 2] = bc.y; o_coords[4 * g + 3] = bc.z;
        o_tidx[g] = idx;
    }
}

/* Brute-force hemisphere ray over all soup triangles: closest t >= mind with
 * the hemisphere predicate (no BVH).  Used to pin AO (BVH-independent). */
ORC_API float orc_hemi_bruteforce(const float *o3, const float *d3_, float mind, const float *spos,
                                  uint32_t nsoup)
{
    v3 o = ld3(o3, 0), d = ld3(d3_, 0);
    float best = FLT_MAX;
    for (uint32_t t = 0; t < nsoup; t += 3) {
        float tt = hemi_tri(o, d, ld4(spos, t), ld4(spos, t + 1), ld4(spos, t + 2));
        if (tt >= mind && tt < best) best = tt;
    }
    return best;
}
