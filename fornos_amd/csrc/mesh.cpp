// mesh.cpp -- seeded procedural meshes and the reference's normal / tangent
// frame computations (Src/mesh.cpp:650-768), plus the fornos_mesh_* ABI.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <random>
#include <stdexcept>
#include <tuple>
#include <unordered_map>

#include "host.h"

using namespace fornos;

namespace {

// Texcoords of the generators span [kUvInset, 1 - kUvInset]: with a UV of
// exactly 0 the reference's raster bound uint32_t(roundf(-0.5)) wraps and the
// triangle is skipped (Src/compute.cpp:136-139), which would leave a column of
// texels unbaked.  2^-12 keeps every bound >= 0 up to 4096^2 maps.
constexpr float kUvInset = 1.0f / 4096.0f;
float uv_inset(float u) { return kUvInset + u * (1.0f - 2.0f * kUvInset); }

// Uniform float in [0,1) from a 32-bit draw.
float unit(std::mt19937& g) { return (float)(g() >> 8) * (1.0f / 16777216.0f); }

// One vertex per triangle corner, as the OBJ loader builds them.
void add_tri(fornos_mesh& m, uint32_t p0, uint32_t p1, uint32_t p2, uint32_t t0, uint32_t t1,
             uint32_t t2, uint32_t n0, uint32_t n1, uint32_t n2)
{
    uint32_t base = m.nv();
    uint32_t v[9] = {p0, t0, n0, p1, t1, n1, p2, t2, n2};
    m.vertices.insert(m.vertices.end(), v, v + 9);
    m.triangles.push_back(base);
    m.triangles.push_back(base + 1);
    m.triangles.push_back(base + 2);
}

// ---- Perlin noise (improved noise, Perlin 2002) with a seeded permutation --
struct Perlin {
    int perm[512];
    explicit Perlin(uint64_t seed)
    {
        std::mt19937 g((uint32_t)(seed * 2654435761u + 17u));
        int p[256];
        for (int i = 0; i < 256; ++i) p[i] = i;
        for (int i = 255; i > 0; --i) {
            int j = (int)(g() % (uint32_t)(i + 1));
            std::swap(p[i], p[j]);
        }
        for (int i = 0; i < 512; ++i) perm[i] = p[i & 255];
    }
    static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
    static double lerp(double t, double a, double b) { return a + t * (b - a); }
    static double grad(int h, double x, double y, double z)
    {
        h &= 15;
        double u = h < 8 ? x : y, v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
        return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
    }
    double noise(double x, double y, double z) const
    {
        double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
        x -= fx; y -= fy; z -= fz;
        double u = fade(x), v = fade(y), w = fade(z);
        int A = perm[X] + Y, AA = perm[A] + Z, AB = perm[A + 1] + Z;
        int B = perm[X + 1] + Y, BA = perm[B] + Z, BB = perm[B + 1] + Z;
        return lerp(w,
                    lerp(v, lerp(u, grad(perm[AA], x, y, z), grad(perm[BA], x - 1, y, z)),
                         lerp(u, grad(perm[AB], x, y - 1, z), grad(perm[BB], x - 1, y - 1, z))),
                    lerp(v, lerp(u, grad(perm[AA + 1], x, y, z - 1), grad(perm[BA + 1], x - 1, y, z - 1)),
                         lerp(u, grad(perm[AB + 1], x, y - 1, z - 1),
                              grad(perm[BB + 1], x - 1, y - 1, z - 1))));
    }
    // fBm: sum_o gain^o * noise(freq0 * lacunarity^o * p)
    double fbm(double x, double y, double z, int octaves, double freq0) const
    {
        double s = 0, amp = 1, f = freq0;
        for (int o = 0; o < octaves; ++o) {
            s += amp * noise(x * f, y * f, z * f);
            amp *= 0.5;
            f *= 2.0;
        }
        return s;
    }
};

constexpr double kPi = 3.14159265358979323846;

void gen_uv_sphere(fornos_mesh& m, float radius, uint32_t segs, uint32_t rings)
{
    // positions: north pole, (rings-1) x (segs+1) ring points (the seam column
    // repeats column 0 bit for bit), south pole.
    m.positions.push_back(mk(0, 0, radius));
    for (uint32_t k = 1; k < rings; ++k) {
        double th = kPi * k / rings;
        for (uint32_t j = 0; j <= segs; ++j) {
            if (j == segs) { m.positions.push_back(m.positions[m.positions.size() - segs]); continue; }
            double ph = 2 * kPi * j / segs;
            m.positions.push_back(mk((float)(radius * std::sin(th) * std::cos(ph)),
                                     (float)(radius * std::sin(th) * std::sin(ph)),
                                     (float)(radius * std::cos(th))));
        }
    }
    m.positions.push_back(mk(0, 0, -radius));
    for (auto& p : m.positions) m.normals.push_back(p * (1.0f / radius));
    auto ring = [&](uint32_t k, uint32_t j) { return 1 + (k - 1) * (segs + 1) + j; };
    const uint32_t south = (uint32_t)m.positions.size() - 1;
    // texcoords: grid points (k = 1..rings-1, j = 0..segs), then pole points per segment
    for (uint32_t k = 1; k < rings; ++k)
        for (uint32_t j = 0; j <= segs; ++j)
            m.texcoords.push_back({uv_inset((float)j / segs), uv_inset((float)k / rings)});
    auto tex = [&](uint32_t k, uint32_t j) { return (k - 1) * (segs + 1) + j; };
    const uint32_t tnorth = (uint32_t)m.texcoords.size();
    for (uint32_t j = 0; j < segs; ++j) m.texcoords.push_back({uv_inset((j + 0.5f) / segs), uv_inset(0.0f)});
    const uint32_t tsouth = (uint32_t)m.texcoords.size();
    for (uint32_t j = 0; j < segs; ++j) m.texcoords.push_back({uv_inset((j + 0.5f) / segs), uv_inset(1.0f)});
    for (uint32_t j = 0; j < segs; ++j)
        add_tri(m, 0, ring(1, j), ring(1, j + 1), tnorth + j, tex(1, j), tex(1, j + 1), 0, ring(1, j),
                ring(1, j + 1));
    for (uint32_t k = 1; k + 1 < rings; ++k)
        for (uint32_t j = 0; j < segs; ++j) {
            uint32_t a = ring(k, j), b = ring(k + 1, j), c = ring(k + 1, j + 1), d = ring(k, j + 1);
            add_tri(m, a, b, c, tex(k, j), tex(k + 1, j), tex(k + 1, j + 1), a, b, c);
            add_tri(m, a, c, d, tex(k, j), tex(k + 1, j + 1), tex(k, j + 1), a, c, d);
        }
    for (uint32_t j = 0; j < segs; ++j) {
        uint32_t a = ring(rings - 1, j), c = ring(rings - 1, j + 1);
        add_tri(m, a, south, c, tex(rings - 1, j), tsouth + j, tex(rings - 1, j + 1), a, south, c);
    }
}

void gen_icosphere_waves(fornos_mesh& m, float radius, uint32_t levels, float amp, uint64_t seed)
{
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<double> P = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                             0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    for (size_t i = 0; i < P.size(); i += 3) {
        double l = std::sqrt(P[i] * P[i] + P[i + 1] * P[i + 1] + P[i + 2] * P[i + 2]);
        P[i] /= l; P[i + 1] /= l; P[i + 2] /= l;
    }
    std::vector<uint32_t> F = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                               11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                               3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (uint32_t lv = 0; lv < levels; ++lv) {
        std::unordered_map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double x = P[3 * a] + P[3 * b], y = P[3 * a + 1] + P[3 * b + 1], z = P[3 * a + 2] + P[3 * b + 2];
            double l = std::sqrt(x * x + y * y + z * z);
            uint32_t id = (uint32_t)(P.size() / 3);
            P.push_back(x / l); P.push_back(y / l); P.push_back(z / l);
            mid[key] = id;
            return id;
        };
        std::vector<uint32_t> G;
        G.reserve(F.size() * 4);
        for (size_t i = 0; i < F.size(); i += 3) {
            uint32_t a = F[i], b = F[i + 1], c = F[i + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t f[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            G.insert(G.end(), f, f + 12);
        }
        F.swap(G);
    }
    std::mt19937 g((uint32_t)seed);
    double dir[8][3], freq[8], phase[8];
    for (int k = 0; k < 8; ++k) {
        double z = 2.0 * unit(g) - 1.0, ph = 2 * kPi * unit(g), s = std::sqrt(std::max(0.0, 1 - z * z));
        dir[k][0] = s * std::cos(ph); dir[k][1] = s * std::sin(ph); dir[k][2] = z;
        freq[k] = 2.0 + 14.0 * unit(g);
        phase[k] = 2 * kPi * unit(g);
    }
    const size_t np = P.size() / 3;
    m.positions.resize(np);
    for (size_t i = 0; i < np; ++i) {
        double h = 0;
        for (int k = 0; k < 8; ++k)
            h += std::sin(freq[k] * (dir[k][0] * P[3 * i] + dir[k][1] * P[3 * i + 1] + dir[k][2] * P[3 * i + 2]) + phase[k]);
        double r = radius + amp * h;
        m.positions[i] = mk((float)(r * P[3 * i]), (float)(r * P[3 * i + 1]), (float)(r * P[3 * i + 2]));
    }
    for (size_t i = 0; i < F.size(); i += 3)
        add_tri(m, F[i], F[i + 1], F[i + 2], UINT32_MAX, UINT32_MAX, UINT32_MAX, F[i], F[i + 1], F[i + 2]);
    vertex_normals(m);
}

struct Heightfield {
    Perlin noise;
    double amp;
    std::vector<double> bx, by, br;
    Heightfield(uint64_t seed, double a, uint32_t bumps) : noise(seed), amp(a)
    {
        std::mt19937 g((uint32_t)(seed ^ 0x9E3779B9u));
        for (uint32_t i = 0; i < bumps; ++i) {
            bx.push_back(unit(g)); by.push_back(unit(g)); br.push_back(0.005 + 0.025 * unit(g));
        }
    }
    double base(double x, double y, int octaves) const { return amp * noise.fbm(x, y, 0.5, octaves, 4.0); }
};

void grid_mesh(fornos_mesh& m, uint32_t n, const std::vector<double>& z)
{
    m.positions.resize((size_t)(n + 1) * (n + 1));
    m.texcoords.resize(m.positions.size());
    for (uint32_t j = 0; j <= n; ++j)
        for (uint32_t i = 0; i <= n; ++i) {
            size_t k = (size_t)j * (n + 1) + i;
            float x = (float)i / n, y = (float)j / n;
            m.positions[k] = mk(x, y, (float)z[k]);
            m.texcoords[k] = {uv_inset(x), uv_inset(y)};
        }
    m.vertices.reserve((size_t)n * n * 18);
    m.triangles.reserve((size_t)n * n * 6);
    for (uint32_t j = 0; j < n; ++j)
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t a = j * (n + 1) + i, b = a + 1, c = a + n + 2, d = a + n + 1;
            add_tri(m, a, b, c, a, b, c, a, b, c);
            add_tri(m, a, c, d, a, c, d, a, c, d);
        }
    vertex_normals(m);
}

void gen_heightfield(fornos_mesh& m, uint32_t n, float amp, uint32_t bumps, uint64_t seed, bool low)
{
    Heightfield hf(seed, amp, bumps);
    std::vector<double> z((size_t)(n + 1) * (n + 1));
    for (uint32_t j = 0; j <= n; ++j)
        for (uint32_t i = 0; i <= n; ++i)
            z[(size_t)j * (n + 1) + i] = hf.base((double)i / n, (double)j / n, low ? 2 : 6);
    if (!low) {
        for (size_t b = 0; b < hf.bx.size(); ++b) {
            double r = hf.br[b];
            int i0 = std::max(0, (int)std::floor((hf.bx[b] - r) * n)), i1 = std::min((int)n, (int)std::ceil((hf.bx[b] + r) * n));
            int j0 = std::max(0, (int)std::floor((hf.by[b] - r) * n)), j1 = std::min((int)n, (int)std::ceil((hf.by[b] + r) * n));
            for (int j = j0; j <= j1; ++j)
                for (int i = i0; i <= i1; ++i) {
                    double dx = (double)i / n - hf.bx[b], dy = (double)j / n - hf.by[b];
                    double q = r * r - dx * dx - dy * dy;
                    if (q > 0) z[(size_t)j * (n + 1) + i] += std::sqrt(q);
                }
        }
    }
    grid_mesh(m, n, z);
}

// (p,q) torus knot tube: centre curve c(u) = ((2+cos qu) cos pu, (2+cos qu) sin pu, -sin qu),
// tube radius R_t, frame from the curve's tangent and its second derivative.
void gen_torus_knot(fornos_mesh& m, uint32_t nu, uint32_t nv, float amp, uint64_t seed)
{
    const double p = 3, q = 5, tube = 0.4;
    Perlin noise(seed);
    auto curve = [&](double u, double* c) {
        double r = 2 + std::cos(q * u);
        c[0] = r * std::cos(p * u); c[1] = r * std::sin(p * u); c[2] = -std::sin(q * u);
    };
    m.positions.resize((size_t)nu * nv);
    for (uint32_t i = 0; i < nu; ++i) {
        double u = 2 * kPi * i / nu, c0[3], c1[3], c2[3], e = 1e-4;
        curve(u, c0); curve(u + e, c1); curve(u - e, c2);
        double T[3], N[3], B[3];
        for (int k = 0; k < 3; ++k) { T[k] = c1[k] - c2[k]; N[k] = c1[k] + c2[k] - 2 * c0[k]; }
        double lt = std::sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]);
        for (int k = 0; k < 3; ++k) T[k] /= lt;
        double dn = N[0] * T[0] + N[1] * T[1] + N[2] * T[2];
        for (int k = 0; k < 3; ++k) N[k] -= dn * T[k];
        double ln = std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
        for (int k = 0; k < 3; ++k) N[k] /= ln;
        B[0] = T[1] * N[2] - T[2] * N[1]; B[1] = T[2] * N[0] - T[0] * N[2]; B[2] = T[0] * N[1] - T[1] * N[0];
        for (uint32_t j = 0; j < nv; ++j) {
            double v = 2 * kPi * j / nv, cv = std::cos(v), sv = std::sin(v);
            double x[3];
            for (int k = 0; k < 3; ++k) x[k] = c0[k] + tube * (cv * N[k] + sv * B[k]);
            double r = tube;
            if (amp != 0) r += amp * noise.fbm(x[0], x[1], x[2], 6, 2.0);
            for (int k = 0; k < 3; ++k) x[k] = c0[k] + r * (cv * N[k] + sv * B[k]);
            m.positions[(size_t)i * nv + j] = mk((float)x[0], (float)x[1], (float)x[2]);
        }
    }
    // cylindrical UVs on a (nu+1) x (nv+1) grid so the seams carry distinct texcoords
    for (uint32_t i = 0; i <= nu; ++i)
        for (uint32_t j = 0; j <= nv; ++j)
            m.texcoords.push_back({uv_inset((float)i / nu), uv_inset((float)j / nv)});
    for (uint32_t i = 0; i < nu; ++i)
        for (uint32_t j = 0; j < nv; ++j) {
            uint32_t i1 = (i + 1) % nu, j1 = (j + 1) % nv;
            uint32_t a = i * nv + j, b = i1 * nv + j, c = i1 * nv + j1, d = i * nv + j1;
            uint32_t ta = i * (nv + 1) + j, tb = (i + 1) * (nv + 1) + j, tc = tb + 1, td = ta + 1;
            // counter-clockwise seen from outside the tube: face normals point outward
            add_tri(m, a, c, b, ta, tc, tb, a, c, b);
            add_tri(m, a, d, c, ta, td, tc, a, d, c);
        }
    vertex_normals(m);
}

}  // namespace

namespace fornos {

// Src/mesh.cpp:650-678.  Normals are indexed by position.
void vertex_normals(fornos_mesh& m)
{
    m.normals.assign(m.positions.size(), mk(0, 0, 0));
    const uint32_t nt = m.ntri();
    for (uint32_t t = 0; t < nt; ++t) {
        uint32_t i0 = m.vertices[3 * m.triangles[3 * t]], i1 = m.vertices[3 * m.triangles[3 * t + 1]],
                 i2 = m.vertices[3 * m.triangles[3 * t + 2]];
        V3 p0 = m.positions[i0], p1 = m.positions[i1], p2 = m.positions[i2];
        V3 n = normalize(cross(p1 - p0, p2 - p0));
        m.normals[i0] = m.normals[i0] + n;
        m.normals[i1] = m.normals[i1] + n;
        m.normals[i2] = m.normals[i2] + n;
    }
    for (auto& n : m.normals) n = normalize(n);
    const uint32_t nv = m.nv();
    for (uint32_t v = 0; v < nv; ++v) m.vertices[3 * v + 2] = m.vertices[3 * v];
}

// Src/mesh.cpp:680-721.  Corners whose positions compare equal (the
// std::map<Vector3> key of :682, ordered by :38-41) share one normal: the sum
// of their face normals in triangle order, normalised.
void vertex_normals_aggressive(fornos_mesh& m)
{
    struct Key {
        float x, y, z;
        bool operator<(const Key& o) const { return std::tie(x, y, z) < std::tie(o.x, o.y, o.z); }
    };
    std::map<Key, uint32_t> index;
    std::vector<V3> sums;
    const uint32_t nt = m.ntri();
    for (uint32_t t = 0; t < nt; ++t) {
        uint32_t v[3] = {m.triangles[3 * t], m.triangles[3 * t + 1], m.triangles[3 * t + 2]};
        V3 p0 = m.pos_of(v[0]), p1 = m.pos_of(v[1]), p2 = m.pos_of(v[2]);
        V3 n = normalize(cross(p1 - p0, p2 - p0));
        V3 ps[3] = {p0, p1, p2};
        for (int k = 0; k < 3; ++k) {
            Key key{ps[k].x, ps[k].y, ps[k].z};
            auto it = index.find(key);
            uint32_t id;
            if (it != index.end()) {
                id = it->second;
                sums[id] = sums[id] + n;
            } else {
                id = (uint32_t)sums.size();
                index.emplace(key, id);
                sums.push_back(n);
            }
            m.vertices[3 * v[k] + 2] = id;
        }
    }
    for (auto& s : sums) s = normalize(s);
    m.normals.swap(sums);
}

// Src/mesh.cpp:724-768.
void tangent_space(fornos_mesh& m)
{
    const uint32_t nv = m.nv(), nt = m.ntri();
    m.tangents.assign(nv, mk(0, 0, 0));
    m.bitangents.assign(nv, mk(0, 0, 0));
    for (uint32_t t = 0; t < nt; ++t) {
        uint32_t a = m.triangles[3 * t], b = m.triangles[3 * t + 1], c = m.triangles[3 * t + 2];
        V3 p0 = m.pos_of(a), p1 = m.pos_of(b), p2 = m.pos_of(c);
        V2 u0 = m.texcoords[m.vertices[3 * a + 1]], u1 = m.texcoords[m.vertices[3 * b + 1]],
           u2 = m.texcoords[m.vertices[3 * c + 1]];
        V3 e0 = p1 - p0, e1 = p2 - p0;
        V2 t0 = {u1.x - u0.x, u1.y - u0.y}, t1 = {u2.x - u0.x, u2.y - u0.y};
        float r = 1.0f / (t0.x * t1.y - t1.x * t0.y);
        V3 sdir = (e0 * t1.y - e1 * t0.y) * r;
        V3 tdir = (e1 * t0.x - e0 * t1.x) * r;
        m.tangents[a] = m.tangents[a] + sdir; m.tangents[b] = m.tangents[b] + sdir;
        m.tangents[c] = m.tangents[c] + sdir;
        m.bitangents[a] = m.bitangents[a] + tdir; m.bitangents[b] = m.bitangents[b] + tdir;
        m.bitangents[c] = m.bitangents[c] + tdir;
    }
    for (uint32_t i = 0; i < nv; ++i) {
        V3 n = m.normals[m.vertices[3 * i + 2]];
        V3 t1 = m.tangents[i];
        m.tangents[i] = normalize(t1 - n * dot(n, t1));
        m.bitangents[i] = cross(n, m.tangents[i]);
    }
}

}  // namespace fornos

extern "C" {

fornos_status fornos_mesh_generate(int generator, uint64_t seed, const fornos_gen_params* p,
                                   fornos_mesh** out)
{
    if (!p || !out) return set_error(FORNOS_INVALID_ARG, "fornos_mesh_generate: null argument");
    *out = nullptr;
    fornos_mesh* m = new (std::nothrow) fornos_mesh();
    if (!m) return set_error(FORNOS_OOM, "fornos_mesh_generate: out of host memory");
    try {
        switch (generator) {
        case FORNOS_GEN_UV_SPHERE:
            if (p->b < 3 || p->c < 2) throw std::invalid_argument("uv sphere needs >= 3 segments and >= 2 rings");
            gen_uv_sphere(*m, p->a, p->b, p->c);
            break;
        case FORNOS_GEN_ICOSPHERE_WAVES:
            if (p->b > 10) throw std::invalid_argument("icosphere subdivisions must be <= 10");
            gen_icosphere_waves(*m, 1.0f, p->b, p->a, seed);
            break;
        case FORNOS_GEN_HEIGHTFIELD:
        case FORNOS_GEN_HEIGHTFIELD_LOW:
            if (p->b < 1 || p->b > 8192) throw std::invalid_argument("heightfield quads must be in [1, 8192]");
            gen_heightfield(*m, p->b, p->a, p->c, seed, generator == FORNOS_GEN_HEIGHTFIELD_LOW);
            break;
        case FORNOS_GEN_TORUS_KNOT:
            if (p->b < 3 || p->c < 3) throw std::invalid_argument("torus knot grid must be >= 3 x 3");
            gen_torus_knot(*m, p->b, p->c, p->a, seed);
            break;
        default:
            throw std::invalid_argument("unknown generator");
        }
    } catch (const std::invalid_argument& e) {
        delete m;
        return set_error(FORNOS_INVALID_ARG, std::string("fornos_mesh_generate: ") + e.what());
    } catch (const std::bad_alloc&) {
        delete m;
        return set_error(FORNOS_OOM, "fornos_mesh_generate: out of host memory");
    }
    *out = m;
    return ok();
}

fornos_status fornos_mesh_from_view(const fornos_mesh_view* v, fornos_mesh** out)
{
    if (!v || !out || !v->positions || !v->vertices || !v->triangles)
        return set_error(FORNOS_INVALID_ARG, "fornos_mesh_from_view: null argument");
    fornos_mesh* m = new fornos_mesh();
    m->positions.resize(v->n_positions);
    std::memcpy(m->positions.data(), v->positions, sizeof(float) * 3 * v->n_positions);
    if (v->normals) {
        m->normals.resize(v->n_normals);
        std::memcpy(m->normals.data(), v->normals, sizeof(float) * 3 * v->n_normals);
    }
    if (v->texcoords) {
        m->texcoords.resize(v->n_texcoords);
        std::memcpy(m->texcoords.data(), v->texcoords, sizeof(float) * 2 * v->n_texcoords);
    }
    m->vertices.assign(v->vertices, v->vertices + 3 * (size_t)v->n_vertices);
    m->triangles.assign(v->triangles, v->triangles + 3 * (size_t)v->n_triangles);
    if (v->tangents && v->bitangents) {
        m->tangents.resize(v->n_vertices);
        m->bitangents.resize(v->n_vertices);
        std::memcpy(m->tangents.data(), v->tangents, sizeof(float) * 3 * v->n_vertices);
        std::memcpy(m->bitangents.data(), v->bitangents, sizeof(float) * 3 * v->n_vertices);
    }
    for (size_t i = 0; i < m->triangles.size(); ++i)
        if (m->triangles[i] >= v->n_vertices) {
            delete m;
            return set_error(FORNOS_INVALID_ARG, "fornos_mesh_from_view: triangle references a missing vertex");
        }
    for (uint32_t i = 0; i < v->n_vertices; ++i)
        if (m->vertices[3 * i] >= v->n_positions ||
            (m->vertices[3 * i + 1] != UINT32_MAX && m->vertices[3 * i + 1] >= v->n_texcoords) ||
            (m->vertices[3 * i + 2] != UINT32_MAX && m->vertices[3 * i + 2] >= v->n_normals)) {
            delete m;
            return set_error(FORNOS_INVALID_ARG, "fornos_mesh_from_view: vertex index out of range");
        }
    *out = m;
    return ok();
}

fornos_status fornos_mesh_copy(const fornos_mesh* m, fornos_mesh** out)
{
    if (!m || !out) return set_error(FORNOS_INVALID_ARG, "fornos_mesh_copy: null argument");
    *out = new fornos_mesh(*m);
    return ok();
}

fornos_status fornos_mesh_view_get(const fornos_mesh* m, fornos_mesh_view* v)
{
    if (!m || !v) return set_error(FORNOS_INVALID_ARG, "fornos_mesh_view_get: null argument");
    v->positions = (const float*)m->positions.data();
    v->n_positions = (uint32_t)m->positions.size();
    v->normals = (const float*)m->normals.data();
    v->n_normals = (uint32_t)m->normals.size();
    v->texcoords = m->texcoords.empty() ? nullptr : (const float*)m->texcoords.data();
    v->n_texcoords = (uint32_t)m->texcoords.size();
    v->tangents = m->tangents.empty() ? nullptr : (const float*)m->tangents.data();
    v->bitangents = m->bitangents.empty() ? nullptr : (const float*)m->bitangents.data();
    v->vertices = m->vertices.data();
    v->n_vertices = m->nv();
    v->triangles = m->triangles.data();
    v->n_triangles = m->ntri();
    return ok();
}

fornos_status fornos_mesh_compute_vertex_normals(fornos_mesh* m)
{
    if (!m) return set_error(FORNOS_INVALID_ARG, "fornos_mesh_compute_vertex_normals: null mesh");
    vertex_normals(*m);
    return ok();
}

fornos_status fornos_mesh_compute_vertex_normals_aggressive(fornos_mesh* m)
{
    if (!m) return set_error(FORNOS_INVALID_ARG, "fornos_mesh_compute_vertex_normals_aggressive: null mesh");
    vertex_normals_aggressive(*m);
    return ok();
}

fornos_status fornos_mesh_compute_tangent_space(fornos_mesh* m)
{
    if (!m) return set_error(FORNOS_INVALID_ARG, "fornos_mesh_compute_tangent_space: null mesh");
    for (uint32_t v = 0; v < m->nv(); ++v)
        if (m->vertices[3 * v + 1] == UINT32_MAX || m->vertices[3 * v + 2] == UINT32_MAX)
            return set_error(FORNOS_INVALID_ARG, "fornos_mesh_compute_tangent_space: mesh lacks texcoords or normals");
    tangent_space(*m);
    return ok();
}

void fornos_mesh_destroy(fornos_mesh* m) { delete m; }

}  // extern "C"
