// meshio.cpp -- Mesh::loadFile (Src/mesh.cpp:235-617): Wavefront OBJ and PLY
// readers producing the reference's Mesh layout, and Mesh::computeFaceNormals
// (Src/mesh.cpp:627-648).
//
// OBJ follows the reference's reader: one statement per line, the first token
// decides (v, vt, vn, f; comments and other statements ignored), numbers read
// with strtod / strtol (a value that does not parse, a range error, or
// trailing text fails the file), face corners v, v/t, v//n or v/t/n with the
// same form on every corner of the file, polygons fanned from their first
// corner, one mesh vertex per face corner (no sharing).  Unlike the
// reference, indices out of range (including the relative negative form,
// which the reference does not support either) are an error instead of
// undefined behaviour.
//
// PLY follows the reference's use of tinyply: vertex x, y, z; optional s, t;
// normals from the vertices (nx, ny, nz) or else from the faces; one mesh
// vertex per face corner, each texcoord / normal index equal to the position
// index (face normals: the face's index).  Read here: ascii and binary
// little / big endian, any numeric property type.  Beyond the reference:
// signed index types (the reference reads only unsigned ones) and polygons
// (fanned; the reference assumes triangles).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "host.h"

using namespace fornos;

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;

bool read_float(const char*& p, float& v, bool required)
{
    char* end;
    errno = 0;
    v = (float)std::strtod(p, &end);
    if (required && p == end) return false;
    p = end;
    return errno == 0;
}

bool read_int(const char*& p, int& v, bool required)
{
    char* end;
    errno = 0;
    v = (int)std::strtol(p, &end, 10);
    if (required && p == end) return false;
    p = end;
    return errno == 0;
}

bool end_of_line(const char* p)
{
    for (; *p; ++p)
        if (!std::isspace((unsigned char)*p)) return false;
    return true;
}

enum class Tok { Unknown, Comment, V, VT, VN, F };

Tok token(const std::string& t)
{
    if (t[0] == '#') return Tok::Comment;
    if (t == "v") return Tok::V;
    if (t == "vt") return Tok::VT;
    if (t == "vn") return Tok::VN;
    if (t == "f") return Tok::F;
    return Tok::Unknown;
}

fornos_status load_obj(const char* path, fornos_mesh& m)
{
    std::ifstream in(path);
    if (!in) return set_error(FORNOS_INVALID_ARG, std::string("cannot open ") + path);
    std::vector<V3> pos, nrm;
    std::vector<V2> tex;
    struct Corner { int v, t, n; };
    std::vector<Corner> corners;
    std::vector<std::pair<size_t, size_t>> faces;  // first corner, corner count
    bool opts_set = false, has_t = false, has_n = false;
    std::string line;
    size_t lineno = 0;
    auto fail = [&](const char* what) {
        return set_error(FORNOS_INVALID_ARG, std::string(path) + ":" + std::to_string(lineno) + ": " + what);
    };
    while (std::getline(in, line)) {
        ++lineno;
        size_t b = 0;
        while (b < line.size() && std::isspace((unsigned char)line[b])) ++b;
        if (b == line.size()) continue;
        size_t e = b;
        while (e < line.size() && !std::isspace((unsigned char)line[e])) ++e;
        const char* p = line.c_str() + e;
        switch (token(line.substr(b, e - b))) {
        case Tok::V: {
            float x, y, z, w;
            if (!read_float(p, x, true) || !read_float(p, y, true) || !read_float(p, z, true) ||
                !read_float(p, w, false) || !end_of_line(p))
                return fail("bad vertex");
            pos.push_back(mk(x, y, z));
        } break;
        case Tok::VT: {
            float u, v, w;
            if (!read_float(p, u, true) || !read_float(p, v, true) || !read_float(p, w, false) || !end_of_line(p))
                return fail("bad texture coordinate");
            tex.push_back(V2{u, v});
        } break;
        case Tok::VN: {
            float i, j, k;
            if (!read_float(p, i, true) || !read_float(p, j, true) || !read_float(p, k, true) || !end_of_line(p))
                return fail("bad normal");
            nrm.push_back(mk(i, j, k));
        } break;
        case Tok::F: {
            const size_t first = corners.size();
            while (!end_of_line(p)) {
                Corner c{0, 0, 0};
                if (!read_int(p, c.v, true)) return fail("bad face vertex");
                if (*p == '/') {
                    ++p;
                    if (!read_int(p, c.t, false)) return fail("bad face texture index");
                }
                if (*p == '/') {
                    ++p;
                    if (!read_int(p, c.n, true)) return fail("bad face normal index");
                }
                if (!opts_set) {
                    has_t = c.t != 0;
                    has_n = c.n != 0;
                    opts_set = true;
                } else if ((c.t != 0) != has_t || (c.n != 0) != has_n) {
                    return fail("face corners mix forms");
                }
                corners.push_back(c);
                while (std::isspace((unsigned char)*p)) ++p;
            }
            faces.push_back({first, corners.size() - first});
        } break;
        default:
            break;
        }
    }
    m = fornos_mesh();
    m.positions = pos;
    m.texcoords = tex;
    m.normals = nrm;
    for (const auto& f : faces) {
        uint32_t v0 = kNone, prev = kNone;
        for (size_t i = 0; i < f.second; ++i) {
            const Corner& c = corners[f.first + i];
            if (c.v < 1 || (size_t)c.v > pos.size() || (c.t && (c.t < 1 || (size_t)c.t > tex.size())) ||
                (c.n && (c.n < 1 || (size_t)c.n > nrm.size())))
                return set_error(FORNOS_INVALID_ARG, std::string(path) + ": face index out of range");
            const uint32_t vi = m.nv();
            m.vertices.push_back((uint32_t)(c.v - 1));
            m.vertices.push_back(c.t ? (uint32_t)(c.t - 1) : kNone);
            m.vertices.push_back(c.n ? (uint32_t)(c.n - 1) : kNone);
            if (v0 == kNone) {
                v0 = vi;
            } else if (prev == kNone) {
                prev = vi;
            } else {
                m.triangles.insert(m.triangles.end(), {v0, prev, vi});
                prev = vi;
            }
        }
    }
    return FORNOS_OK;
}

// ---- PLY -------------------------------------------------------------------
struct PlyProp {
    std::string name, type, count_type;  // count_type empty: scalar
};
struct PlyElem {
    std::string name;
    size_t count = 0;
    std::vector<PlyProp> props;
};

size_t type_size(const std::string& t)
{
    if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
    if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
    if (t == "int" || t == "uint" || t == "int32" || t == "uint32" || t == "float" || t == "float32") return 4;
    if (t == "double" || t == "float64") return 8;
    return 0;
}

struct PlyReader {
    std::ifstream& in;
    int format;  // 0 ascii, 1 little endian, 2 big endian
    std::istringstream line;
    bool next_ascii(std::string& tok)
    {
        while (!(line >> tok)) {
            std::string l;
            if (!std::getline(in, l)) return false;
            line.clear();
            line.str(l);
        }
        return true;
    }
    bool value(const std::string& t, double& out)
    {
        if (format == 0) {
            std::string tok;
            if (!next_ascii(tok)) return false;
            char* end;
            out = std::strtod(tok.c_str(), &end);
            return end != tok.c_str();
        }
        const size_t n = type_size(t);
        unsigned char b[8];
        if (!n || !in.read((char*)b, n)) return false;
        if (format == 2) std::reverse(b, b + n);
        if (t == "char" || t == "int8") out = (int8_t)b[0];
        else if (t == "uchar" || t == "uint8") out = b[0];
        else if (t == "short" || t == "int16") { int16_t v; std::memcpy(&v, b, 2); out = v; }
        else if (t == "ushort" || t == "uint16") { uint16_t v; std::memcpy(&v, b, 2); out = v; }
        else if (t == "int" || t == "int32") { int32_t v; std::memcpy(&v, b, 4); out = v; }
        else if (t == "uint" || t == "uint32") { uint32_t v; std::memcpy(&v, b, 4); out = v; }
        else if (t == "float" || t == "float32") { float v; std::memcpy(&v, b, 4); out = v; }
        else { double v; std::memcpy(&v, b, 8); out = v; }
        return true;
    }
};

fornos_status load_ply(const char* path, fornos_mesh& m)
{
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) return set_error(FORNOS_INVALID_ARG, std::string("cannot open ") + path);
    // every list entry takes at least one byte of the file: a list count
    // above the file's size is corrupt (and must not size an allocation)
    const double file_bytes = (double)in.tellg();
    in.seekg(0);
    auto bad = [&](const std::string& what) { return set_error(FORNOS_INVALID_ARG, std::string(path) + ": " + what); };
    std::string l;
    if (!std::getline(in, l) || l.compare(0, 3, "ply") != 0) return bad("not a PLY file");
    int format = -1;
    std::vector<PlyElem> elems;
    while (std::getline(in, l)) {
        if (!l.empty() && l.back() == '\r') l.pop_back();
        std::istringstream s(l);
        std::string kw;
        s >> kw;
        if (kw == "format") {
            std::string f;
            s >> f;
            format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
        } else if (kw == "element") {
            PlyElem e;
            s >> e.name >> e.count;
            elems.push_back(e);
        } else if (kw == "property") {
            if (elems.empty()) return bad("property before element");
            PlyProp p;
            s >> p.type;
            if (p.type == "list") {
                s >> p.count_type >> p.type;
            }
            s >> p.name;
            if (!type_size(p.type) || (!p.count_type.empty() && !type_size(p.count_type)))
                return bad("unknown property type");
            elems.back().props.push_back(p);
        } else if (kw == "end_header") {
            break;
        }
    }
    if (format < 0) return bad("unsupported format");
    PlyReader r{in, format, {}};
    std::vector<double> vx, vy, vz, vs, vt, vnx, vny, vnz, fnx, fny, fnz;
    std::vector<std::vector<uint32_t>> faces;
    bool has_st = false, has_vn = false, has_fn = false;
    for (const PlyElem& e : elems) {
        const bool is_v = e.name == "vertex", is_f = e.name == "face";
        auto has = [&](const char* n) {
            for (const PlyProp& p : e.props)
                if (p.name == n && p.count_type.empty()) return true;
            return false;
        };
        if (is_v) {
            if (!has("x") || !has("y") || !has("z")) return bad("vertex element without x, y, z");
            has_st = has("s") && has("t");
            has_vn = has("nx") && has("ny") && has("nz");
        }
        if (is_f) has_fn = has("nx") && has("ny") && has("nz");
        for (size_t i = 0; i < e.count; ++i) {
            for (const PlyProp& p : e.props) {
                if (!p.count_type.empty()) {
                    double cnt;
                    if (!r.value(p.count_type, cnt)) return bad("truncated list");
                    if (!(cnt >= 0 && cnt <= file_bytes) || cnt != std::floor(cnt))
                        return bad("bad list count");
                    std::vector<uint32_t> idx((size_t)cnt);
                    for (size_t k = 0; k < idx.size(); ++k) {
                        double v;
                        if (!r.value(p.type, v)) return bad("truncated list");
                        if (v < 0 || v >= 4294967295.0) return bad("negative or huge vertex index");
                        idx[k] = (uint32_t)v;
                    }
                    if (is_f && (p.name == "vertex_indices" || p.name == "vertex_index")) faces.push_back(idx);
                    continue;
                }
                double v;
                if (!r.value(p.type, v)) return bad("truncated data");
                if (is_v) {
                    if (p.name == "x") vx.push_back(v);
                    else if (p.name == "y") vy.push_back(v);
                    else if (p.name == "z") vz.push_back(v);
                    else if (p.name == "s") vs.push_back(v);
                    else if (p.name == "t") vt.push_back(v);
                    else if (p.name == "nx") vnx.push_back(v);
                    else if (p.name == "ny") vny.push_back(v);
                    else if (p.name == "nz") vnz.push_back(v);
                } else if (is_f) {
                    if (p.name == "nx") fnx.push_back(v);
                    else if (p.name == "ny") fny.push_back(v);
                    else if (p.name == "nz") fnz.push_back(v);
                }
            }
        }
    }
    m = fornos_mesh();
    for (size_t i = 0; i < vx.size(); ++i) m.positions.push_back(mk((float)vx[i], (float)vy[i], (float)vz[i]));
    if (has_st)
        for (size_t i = 0; i < vs.size(); ++i) m.texcoords.push_back(V2{(float)vs[i], (float)vt[i]});
    const int nfrom = has_vn ? 1 : has_fn ? 2 : 0;
    if (nfrom == 1)
        for (size_t i = 0; i < vnx.size(); ++i) m.normals.push_back(mk((float)vnx[i], (float)vny[i], (float)vnz[i]));
    if (nfrom == 2)
        for (size_t i = 0; i < fnx.size(); ++i) m.normals.push_back(mk((float)fnx[i], (float)fny[i], (float)fnz[i]));
    for (size_t fi = 0; fi < faces.size(); ++fi) {
        const auto& f = faces[fi];
        if (f.size() < 3) continue;
        for (uint32_t v : f)
            if (v >= m.positions.size()) return bad("face index out of range");
        // fan from the first corner, three new vertices per triangle
        for (size_t k = 1; k + 1 < f.size(); ++k) {
            const uint32_t vstart = m.nv();
            for (uint32_t v : {f[0], f[k], f[k + 1]}) {
                m.vertices.push_back(v);
                m.vertices.push_back(has_st ? v : kNone);
                m.vertices.push_back(nfrom == 1 ? v : nfrom == 2 ? (uint32_t)fi : kNone);
            }
            m.triangles.insert(m.triangles.end(), {vstart, vstart + 1, vstart + 2});
        }
    }
    return FORNOS_OK;
}

bool ends_with(const std::string& s, const char* e)
{
    const size_t n = std::strlen(e);
    return s.size() >= n && s.compare(s.size() - n, n, e) == 0;
}

}  // namespace

namespace fornos {

// Src/mesh.cpp:627-648: one normal per triangle, normalize(cross(p1 - p0, p2 - p0)),
// every corner of the triangle pointing at it.
void face_normals(fornos_mesh& m)
{
    m.normals.clear();
    const uint32_t nt = m.ntri();
    for (uint32_t t = 0; t < nt; ++t) {
        const uint32_t a = m.triangles[3 * t], b = m.triangles[3 * t + 1], c = m.triangles[3 * t + 2];
        const V3 p0 = m.pos_of(a), p1 = m.pos_of(b), p2 = m.pos_of(c);
        const uint32_t ni = (uint32_t)m.normals.size();
        m.normals.push_back(normalize(cross(p1 - p0, p2 - p0)));
        m.vertices[3 * a + 2] = ni;
        m.vertices[3 * b + 2] = ni;
        m.vertices[3 * c + 2] = ni;
    }
}

}  // namespace fornos

extern "C" {

fornos_status fornos_mesh_load(const char* path, fornos_mesh** out)
{
    if (!path || !out) return set_error(FORNOS_INVALID_ARG, "fornos_mesh_load: null argument");
    *out = nullptr;
    try {
        std::unique_ptr<fornos_mesh> m(new fornos_mesh());
        const std::string p(path);
        fornos_status st;
        if (ends_with(p, ".obj")) st = load_obj(path, *m);
        else if (ends_with(p, ".ply")) st = load_ply(path, *m);
        else return set_error(FORNOS_INVALID_ARG, std::string("unsupported mesh file: ") + path);
        if (st) return st;
        *out = m.release();
        return FORNOS_OK;
    } catch (const std::bad_alloc&) {
        return set_error(FORNOS_OOM, std::string("fornos_mesh_load: out of memory reading ") + path);
    } catch (const std::exception& e) {
        return set_error(FORNOS_INVALID_ARG, std::string("fornos_mesh_load: ") + path + ": " + e.what());
    }
}

fornos_status fornos_mesh_compute_face_normals(fornos_mesh* m)
{
    if (!m) return set_error(FORNOS_INVALID_ARG, "fornos_mesh_compute_face_normals: null mesh");
    face_normals(*m);
    return FORNOS_OK;
}

}  // extern "C"
