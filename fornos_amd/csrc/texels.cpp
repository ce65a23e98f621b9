// texels.cpp -- the texel set: MapUV raster (Src/compute.cpp:75-193) and
// CompressedMapUV (Src/compute.cpp:372-416), the sample table
// (Src/math.h:324-382) and the library's error state.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "host.h"
#include "mapuv.h"

using namespace fornos;

namespace fornos {

static thread_local std::string g_last_error;

fornos_status set_error(fornos_status st, const std::string& msg)
{
    g_last_error = msg;
    return st;
}
fornos_status ok() { return FORNOS_OK; }

namespace {

V3 blend(V3 a, V3 b, V3 c, V3 w) { return a * w.x + b * w.y + c * w.z; }

}  // namespace

const char* const kMissingAttributes = "Low poly mesh is missing texture coordinates or normals information";
const char* const kTangentIndex = "Hybrid mapping: a normal index is out of the tangent frame's range";
const char* const kUvBoxLimitMsg = "texture coordinates too far outside [0, 1] for the Hybrid mapping";

// Src/compute.cpp:75-193 (edge <= 0: createMapUV, the Smooth and
// LowPolyNormals methods) or Src/compute.cpp:213-345 (edge > 0:
// createMapUVEdge, the Hybrid method: bounds clamped at 0 only, tangents
// indexed by normal index, direction blended toward the surface normal by
// min(triangleDistance / edge, 1)).
const char* build_texels(const fornos_mesh& lo, const fornos_mesh* dirs, uint32_t W, uint32_t H,
                         fornos_texels& out, float edge)
{
    const bool hybrid = edge > 0.0f;
    const size_t npx = (size_t)W * H;
    out.width = W;
    out.height = H;
    out.map_pos.assign(3 * npx, 0.0f);
    out.map_dir.assign(3 * npx, 0.0f);
    out.map_nrm.assign(3 * npx, 0.0f);
    out.map_tan.assign(3 * npx, 0.0f);
    out.map_bit.assign(3 * npx, 0.0f);
    const bool has_tan = !lo.tangents.empty();
    const uint32_t nt = lo.ntri();
    for (uint32_t t = 0; t < nt; ++t) {
        uint32_t vi[3] = {lo.triangles[3 * t], lo.triangles[3 * t + 1], lo.triangles[3 * t + 2]};
        for (int k = 0; k < 3; ++k)
            if (lo.vertices[3 * vi[k] + 1] == UINT32_MAX || lo.vertices[3 * vi[k] + 2] == UINT32_MAX)
                return kMissingAttributes;
        V3 p[3], n[3], d[3], tg[3], bt[3];
        V2 u[3];
        for (int k = 0; k < 3; ++k) {
            p[k] = lo.positions[lo.vertices[3 * vi[k]]];
            u[k] = lo.texcoords[lo.vertices[3 * vi[k] + 1]];
            n[k] = lo.normals[lo.vertices[3 * vi[k] + 2]];
            d[k] = dirs ? dirs->normals[dirs->vertices[3 * vi[k] + 2]] : n[k];
            const uint32_t ti = hybrid ? lo.vertices[3 * vi[k] + 2] : vi[k];
            if (has_tan && ti >= lo.tangents.size()) return kTangentIndex;
            tg[k] = has_tan ? lo.tangents[ti] : mk(0, 0, 0);
            bt[k] = has_tan ? lo.bitangents[ti] : mk(0, 0, 0);
        }
        const TexelBox bx = texel_box(u, W, H, hybrid);
        if (!bx.ok) return kUvBoxLimitMsg;
        for (uint32_t y = bx.yMin; y <= bx.yMax; ++y)
            for (uint32_t x = bx.xMin; x <= bx.xMax; ++x) {
                V3 b;
                if (texel_inside(u, x, y, W, H, b)) {
                    const size_t i = (size_t)y * W + x;
                    if (i >= npx) continue;  // the reference writes past the end here (y == H)
                    const V3 pp = blend(p[0], p[1], p[2], b), nn = normalize(blend(n[0], n[1], n[2], b));
                    V3 dd = normalize(blend(d[0], d[1], d[2], b));
                    if (hybrid) {
                        dd = hybrid_direction(p[0], p[1], p[2], pp, nn, dd, edge);
                    }
                    st3(out.map_pos.data(), i, pp);
                    st3(out.map_dir.data(), i, dd);
                    st3(out.map_nrm.data(), i, nn);
                    st3(out.map_tan.data(), i, normalize(blend(tg[0], tg[1], tg[2], b)));
                    st3(out.map_bit.data(), i, normalize(blend(bt[0], bt[1], bt[2], b)));
                }
            }
    }
    // CompressedMapUV: texels with dot(dir, dir) > 0.5, row-major.
    out.indices.clear();
    for (size_t i = 0; i < npx; ++i) {
        V3 n = ld3(out.map_dir.data(), i);
        if (dot(n, n) > 0.5f) out.indices.push_back((uint32_t)i);
    }
    const size_t c = out.indices.size();
    out.pos.resize(3 * c); out.dir.resize(3 * c); out.nrm.resize(3 * c);
    out.tan.resize(3 * c); out.bit.resize(3 * c);
    for (size_t k = 0; k < c; ++k) {
        const size_t i = out.indices[k];
        std::memcpy(&out.pos[3 * k], &out.map_pos[3 * i], 12);
        std::memcpy(&out.dir[3 * k], &out.map_dir[3 * i], 12);
        std::memcpy(&out.nrm[3 * k], &out.map_nrm[3 * i], 12);
        std::memcpy(&out.tan[3 * k], &out.map_tan[3 * i], 12);
        std::memcpy(&out.bit[3 * k], &out.map_bit[3 * i], 12);
    }
    return nullptr;
}

}  // namespace fornos

namespace {

// Src/math.h:324-332
uint32_t reverse_bits(uint32_t bits)
{
    bits = (bits << 16u) | (bits >> 16u);
    bits = ((bits & 0x55555555u) << 1u) | ((bits & 0xAAAAAAAAu) >> 1u);
    bits = ((bits & 0x33333333u) << 2u) | ((bits & 0xCCCCCCCCu) >> 2u);
    bits = ((bits & 0x0F0F0F0Fu) << 4u) | ((bits & 0xF0F0F0F0u) >> 4u);
    bits = ((bits & 0x00FF00FFu) << 8u) | ((bits & 0xFF00FF00u) >> 8u);
    return bits;
}

}  // namespace

extern "C" {

const char* fornos_last_error(void) { return g_last_error.c_str(); }

fornos_status fornos_texels_create_hybrid(const fornos_mesh* lo, const fornos_mesh* dirs, uint32_t W, uint32_t H,
                                          float edge, fornos_texels** out)
{
    if (!lo || !out) return set_error(FORNOS_INVALID_ARG, "fornos_texels_create: null argument");
    if (W == 0 || H == 0 || W > 16384 || H > 16384)
        return set_error(FORNOS_INVALID_ARG, "fornos_texels_create: map size must be in [1, 16384]");
    if (dirs && (dirs->nv() != lo->nv()))
        return set_error(FORNOS_INVALID_ARG, "fornos_texels_create: mapping mesh must be a copy of the low-poly mesh");
    if (!(edge >= 0.0f)) return set_error(FORNOS_INVALID_ARG, "fornos_texels_create_hybrid: edge must be >= 0");
    *out = nullptr;
    fornos_texels* t = new (std::nothrow) fornos_texels();
    if (!t) return set_error(FORNOS_OOM, "fornos_texels_create: out of host memory");
    try {
        if (const char* err = build_texels(*lo, dirs, W, H, *t, edge)) {
            delete t;
            return set_error(FORNOS_INVALID_ARG, err);
        }
    } catch (const std::bad_alloc&) {
        delete t;
        return set_error(FORNOS_OOM, "fornos_texels_create: out of host memory");
    }
    *out = t;
    return ok();
}

fornos_status fornos_texels_create(const fornos_mesh* lo, const fornos_mesh* dirs, uint32_t W,
                                   uint32_t H, fornos_texels** out)
{
    return fornos_texels_create_hybrid(lo, dirs, W, H, 0.0f, out);
}

fornos_status fornos_texels_view_get(const fornos_texels* t, fornos_texel_view* v)
{
    if (!t || !v) return set_error(FORNOS_INVALID_ARG, "fornos_texels_view_get: null argument");
    v->count = (uint32_t)t->indices.size();
    v->width = t->width;
    v->height = t->height;
    v->positions = t->pos.data();
    v->directions = t->dir.data();
    v->normals = t->nrm.data();
    v->tangents = t->tan.data();
    v->bitangents = t->bit.data();
    v->indices = t->indices.data();
    return ok();
}

fornos_status fornos_texels_export_mapuv(const fornos_texels* t, float* p, float* d, float* n,
                                         float* tg, float* b)
{
    if (!t) return set_error(FORNOS_INVALID_ARG, "fornos_texels_export_mapuv: null texels");
    const size_t bytes = t->map_pos.size() * sizeof(float);
    if (p) std::memcpy(p, t->map_pos.data(), bytes);
    if (d) std::memcpy(d, t->map_dir.data(), bytes);
    if (n) std::memcpy(n, t->map_nrm.data(), bytes);
    if (tg) std::memcpy(tg, t->map_tan.data(), bytes);
    if (b) std::memcpy(b, t->map_bit.data(), bytes);
    return ok();
}

void fornos_texels_destroy(fornos_texels* t) { delete t; }

// Src/math.h:339-344 (hammersley with Cranley-Patterson shift on x and XOR
// scrambling on y) and :363-382 (cosine-weighted hemisphere directions).
fornos_status fornos_sample_table(uint32_t S, uint32_t P, uint32_t seed, float* out, uint32_t* rand_out)
{
    if (!out || S == 0 || P == 0) return set_error(FORNOS_INVALID_ARG, "fornos_sample_table: bad argument");
    // the permutations' random pairs in draw order, then the samples of each
    // permutation (independent of one another) on several threads
    std::mt19937 re(seed);
    std::vector<uint32_t> rnd(2 * (size_t)P);
    for (uint32_t j = 0; j < P; ++j) {
        rnd[2 * j] = (uint32_t)re();
        rnd[2 * j + 1] = (uint32_t)re();
    }
    if (rand_out) std::memcpy(rand_out, rnd.data(), sizeof(uint32_t) * rnd.size());
    parallel_chunks(P, [&](size_t j0, size_t j1) {
        for (uint32_t j = (uint32_t)j0; j < (uint32_t)j1; ++j) {
            const uint32_t rx = rnd[2 * j], ry = rnd[2 * j + 1];
            for (uint32_t i = 0; i < S; ++i) {
                const float x = (float)i / (float)S + (float)(rx & 0xffffu) / (float)(1 << 16);
                const float y = (float)(reverse_bits(i) ^ ry) * 2.3283064365386963e-10f;
                const float ux = x - (float)(long)x;
                const float r = std::sqrt(ux);
                const float phi = (float)(2.0 * 3.14159265) * y;
                float* o = &out[3 * ((size_t)j * S + i)];
                o[0] = r * std::cos(phi);
                o[1] = r * std::sin(phi);
                o[2] = std::sqrt(1.0f - ux);
            }
        }
    });
    return ok();
}

}  // extern "C"
