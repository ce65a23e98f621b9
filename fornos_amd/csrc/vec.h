// vec.h -- fp32 vector helpers shared by the host C++ and the HIP kernels.
//
// Every operation spells out its evaluation order so that host and device
// produce bit-identical results when both are compiled with
// -ffp-contract=off (no a*b+c contraction) and IEEE-correct division and
// square root (the HIP default for fp32 on gfx950):
//   dot(a,b)     = (a.x*b.x + a.y*b.y) + a.z*b.z          Src/math.h:92-95
//   cross(a,b)                                            Src/math.h:97-103
//   normalize(v) = v * (1.0f / sqrtf(dot(v,v)))           Src/math.h:109-112
//   hmin/hmax    = fminf/fmaxf per component              Src/math.h:114-128
//   gmin/gmax    = GLSL min/max: (y < x) ? y : x / (x < y) ? y : x
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FN __host__ __device__ inline
#else
#define FN inline
#endif

#pragma clang fp contract(off)

namespace fornos {

struct V3 {
    float x, y, z;
};
struct V2 {
    float x, y;
};

FN V3 mk(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
FN V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
FN V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
FN V3 operator-(V3 a) { return mk(-a.x, -a.y, -a.z); }
FN V3 operator*(V3 a, float k) { return mk(a.x * k, a.y * k, a.z * k); }
FN V3 operator/(V3 a, float k) { return mk(a.x / k, a.y / k, a.z / k); }
FN float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
FN V3 cross(V3 a, V3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
FN V3 normalize(V3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
FN V3 hmin(V3 a, V3 b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
FN V3 hmax(V3 a, V3 b) { return mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
FN float gmin(float x, float y) { return (y < x) ? y : x; }
FN float gmax(float x, float y) { return (x < y) ? y : x; }

// edgeDistance / triangleDistance (Src/compute.cpp:195-210): distance of p
// to the line through e0, e1; the smallest over the edges 01, 02, 12.
FN float edge_distance(V3 e0, V3 e1, V3 p)
{
    const V3 v = p - e0, e = e1 - e0;
    const float t = dot(v, e) / dot(e, e);
    const V3 w = (e0 + e * t) - p;
    return sqrtf(dot(w, w));
}
FN float triangle_distance(V3 p0, V3 p1, V3 p2, V3 p)
{
    return fminf(edge_distance(p0, p1, p), fminf(edge_distance(p0, p2, p), edge_distance(p1, p2, p)));
}

FN V3 ld3(const float* p, size_t i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
FN void st3(float* p, size_t i, V3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

}  // namespace fornos
