// mapuv.h -- the MapUV raster arithmetic (Src/compute.cpp:75-349) shared by
// the host raster (texels.cpp) and the device raster (raster.h): the
// barycentric test, the texel bounding box of a triangle in both of the
// reference's forms, and the Hybrid method's direction blend.
#pragma once
#include "vec.h"

#pragma clang fp contract(off)

namespace fornos {

// Src/math.h:250-258
FN V3 bary2(V2 p, V2 p0, V2 p1, V2 p2)
{
    const V2 v0 = {p1.x - p0.x, p1.y - p0.y}, v1 = {p2.x - p0.x, p2.y - p0.y}, v2 = {p.x - p0.x, p.y - p0.y};
    const float s = 1.0f / (v0.x * v1.y - v1.x * v0.y);
    const float i = (v2.x * v1.y - v1.x * v2.y) * s;
    const float j = (v0.x * v2.y - v2.x * v0.y) * s;
    return mk(1.0f - i - j, i, j);
}

// uint32_t(std::roundf(x)) as the reference's x86-64 build evaluates it: a
// 64-bit truncating conversion of which the low 32 bits are kept.
FN uint32_t round_u32(float x) { return (uint32_t)(int64_t)roundf(x); }
FN uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

// Candidate texels of one triangle that a run may visit.  The reference's
// Hybrid raster clamps its bounding box at 0 only; texels beyond the map's
// last row or index are written past the end there and skipped here.
constexpr uint64_t kUvBoxLimit = 1ull << 30;

struct TexelBox {
    uint32_t xMin, yMin, xMax, yMax;
    bool ok;  // false: more than kUvBoxLimit candidates (Hybrid only)
};

// Texel bounding box of a triangle with texture coordinates u[3].
// createMapUV (Src/compute.cpp:132-139) clamps each bound to W / H;
// createMapUVEdge (Src/compute.cpp:272-279) clamps at 0 only.  The Hybrid
// box is then cut to the texels that can land inside the map (rows < H,
// y * W + x < W * H), which leaves the visited in-map texels and their
// order unchanged.
FN TexelBox texel_box(const V2* u, uint32_t W, uint32_t H, bool hybrid)
{
    const V2 scale = {(float)W, (float)H};
    const V2 pixsize = {1.0f / scale.x, 1.0f / scale.y};
    const V2 halfpix = {pixsize.x * 0.5f, pixsize.y * 0.5f};
    V2 s[3];
    for (int k = 0; k < 3; ++k) s[k] = {(u[k].x - halfpix.x) * scale.x, (u[k].y - halfpix.y) * scale.y};
    const float lx = fminf(s[0].x, fminf(s[1].x, s[2].x)), ly = fminf(s[0].y, fminf(s[1].y, s[2].y));
    const float hx = fmaxf(s[0].x, fmaxf(s[1].x, s[2].x)), hy = fmaxf(s[0].y, fmaxf(s[1].y, s[2].y));
    TexelBox b;
    b.ok = true;
    if (!hybrid) {
        b.xMin = umin(round_u32(lx), W);
        b.yMin = umin(round_u32(ly), H);
        b.xMax = umin(round_u32(hx), W);
        b.yMax = umin(round_u32(hy), H);
        return b;
    }
    b.xMin = (uint32_t)(int64_t)fmaxf(roundf(lx), 0.0f);
    b.yMin = (uint32_t)(int64_t)fmaxf(roundf(ly), 0.0f);
    b.xMax = (uint32_t)(int64_t)fmaxf(roundf(hx), 0.0f);
    b.yMax = (uint32_t)(int64_t)fmaxf(roundf(hy), 0.0f);
    if (b.yMax > H - 1) b.yMax = H - 1;
    if (b.yMin > b.yMax || b.xMin > b.xMax) return b;
    const uint64_t last = (uint64_t)W * H - 1 - (uint64_t)b.yMin * W;
    if (b.xMax > last) b.xMax = (uint32_t)last;
    if (b.xMin > b.xMax) return b;
    b.ok = (uint64_t)(b.xMax - b.xMin + 1) * (uint64_t)(b.yMax - b.yMin + 1) <= kUvBoxLimit;
    return b;
}

// The barycentric test of both rasters (Src/compute.cpp:144-150).
FN bool texel_inside(const V2* u, uint32_t x, uint32_t y, uint32_t W, uint32_t H, V3& b)
{
    const V2 pixsize = {1.0f / (float)W, 1.0f / (float)H};
    const V2 halfpix = {pixsize.x * 0.5f, pixsize.y * 0.5f};
    const V2 q = {(float)x * pixsize.x + halfpix.x, (float)y * pixsize.y + halfpix.y};
    b = bary2(q, u[0], u[1], u[2]);
    return b.x >= -0.001f && b.x <= 1 && b.y >= -0.001f && b.y <= 1 && b.z >= -0.001f && b.z <= 1;
}

// Src/compute.cpp:305-307: t = min(triangleDistance / edge, 1) and the ray
// direction normalize(dsmooth * (1 - t) + n * t).
FN V3 hybrid_direction(V3 p0, V3 p1, V3 p2, V3 p, V3 n, V3 dsmooth, float edge)
{
    const float t = fminf(triangle_distance(p0, p1, p2, p) / edge, 1.0f);
    return normalize(dsmooth * (1.0f - t) + n * t);
}

}  // namespace fornos
