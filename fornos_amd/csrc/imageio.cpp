// imageio.cpp -- image files of the export (Src/image.cpp:185-404): PNG and
// TGA for 8-bit RGB, OpenEXR (uncompressed scanlines, half channels) for
// float data, and the exportFloatImage / exportNormalImage /
// exportVectorImage entry points that choose by the file extension.
//
// The reference writes PNG / TGA through stb_image_write and EXR through
// tinyexr.  Encoded bytes differ between encoders (deflate choices, RLE), so
// what is kept is the decoded content: the same RGB pixels in the same
// orientation, and EXR channels named and ordered as the reference names them
// holding the same half values (tinyexr's float -> half rounding, below).
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fornos_export.h"
#include "host.h"

using namespace fornos;

namespace {

bool ends_with(const std::string& s, const char* e)
{
    const size_t n = std::strlen(e);
    return s.size() >= n && s.compare(s.size() - n, n, e) == 0;
}

enum class Ext { Unknown, Png, Tga, Exr };

// getExtension (Src/image.cpp:58-64)
Ext extension(const char* path)
{
    const std::string p(path);
    if (ends_with(p, ".png")) return Ext::Png;
    if (ends_with(p, ".tga")) return Ext::Tga;
    if (ends_with(p, ".exr")) return Ext::Exr;
    return Ext::Unknown;
}

struct File {
    FILE* f = nullptr;
    explicit File(const char* path) : f(std::fopen(path, "wb")) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
    bool put(const void* p, size_t n) { return std::fwrite(p, 1, n, f) == n; }
};

void be32(std::vector<uint8_t>& v, uint32_t x)
{
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void le(std::vector<uint8_t>& v, uint64_t x, int bytes)
{
    for (int i = 0; i < bytes; ++i) v.push_back((uint8_t)(x >> (8 * i)));
}

void png_chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data)
{
    be32(out, (uint32_t)data.size());
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    be32(out, (uint32_t)crc32(0L, out.data() + at, (uInt)(data.size() + 4)));
}

// PNG image data as one zlib stream, built in parallel: the rows are split
// into segments of about 256 KiB, each worker filters its rows (PNG filter 1,
// "Sub": byte minus the byte 3 to the left, mod 256) and deflates them as a
// raw deflate run of its own ending on a byte-aligned sync-flush block (the
// last segment ends the stream).  A run never refers back into another
// segment, so the concatenation is one valid deflate stream; its Adler-32
// is combined from the segments' (adler32_combine).
bool png_idat(const uint8_t* rgb, uint32_t width, uint32_t height, int level, std::vector<uint8_t>& z)
{
    const size_t stride = 3 * (size_t)width, line = stride + 1;
    const uint32_t rows_per = (uint32_t)std::max<size_t>(1, ((size_t)1 << 18) / line);
    const uint32_t nseg = (height + rows_per - 1) / rows_per;
    std::vector<std::vector<uint8_t>> part(nseg);
    std::vector<uLong> adl(nseg);
    std::vector<size_t> len(nseg);
    std::atomic<uint32_t> next{0};
    std::atomic<bool> ok{true};
    auto work = [&]() {
        std::vector<uint8_t> raw;
        for (uint32_t i; (i = next.fetch_add(1)) < nseg && ok.load();) {
            const uint32_t y0 = i * rows_per, y1 = std::min(height, y0 + rows_per);
            raw.resize((size_t)(y1 - y0) * line);
            for (uint32_t y = y0; y < y1; ++y) {
                const uint8_t* r = rgb + (size_t)y * stride;
                uint8_t* o = raw.data() + (size_t)(y - y0) * line;
                o[0] = 1;
                for (size_t k = 0; k < stride; ++k) o[1 + k] = (uint8_t)(r[k] - (k >= 3 ? r[k - 3] : 0));
            }
            len[i] = raw.size();
            adl[i] = adler32(adler32(0L, Z_NULL, 0), raw.data(), (uInt)raw.size());
            z_stream zs{};
            if (deflateInit2(&zs, level, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) {
                ok = false;
                return;
            }
            part[i].resize(deflateBound(&zs, (uLong)raw.size()) + 64);
            zs.next_in = raw.data();
            zs.avail_in = (uInt)raw.size();
            zs.next_out = part[i].data();
            zs.avail_out = (uInt)part[i].size();
            const bool last = i + 1 == nseg;
            const int r = deflate(&zs, last ? Z_FINISH : Z_SYNC_FLUSH);
            if ((last ? r != Z_STREAM_END : r != Z_OK) || zs.avail_in != 0) ok = false;
            part[i].resize(part[i].size() - zs.avail_out);
            deflateEnd(&zs);
        }
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nthreads = std::min<unsigned>(std::min<unsigned>(hw, 16u), nseg);
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nthreads; ++t) pool.emplace_back(work);
    work();
    for (std::thread& t : pool) t.join();
    if (!ok) return false;
    z.assign({0x78, 0x9c});  // deflate, 32 KiB window, no dictionary
    uLong a = adl[0];
    for (uint32_t i = 0; i < nseg; ++i) {
        z.insert(z.end(), part[i].begin(), part[i].end());
        if (i) a = adler32_combine(a, adl[i], (z_off_t)len[i]);
    }
    be32(z, (uint32_t)a);
    return true;
}

// EXR attribute: name\0 type\0 size(int32) value
void exr_attr(std::vector<uint8_t>& v, const char* name, const char* type, const std::vector<uint8_t>& val)
{
    v.insert(v.end(), name, name + std::strlen(name) + 1);
    v.insert(v.end(), type, type + std::strlen(type) + 1);
    le(v, val.size(), 4);
    v.insert(v.end(), val.begin(), val.end());
}

}  // namespace

extern "C" {

// tinyexr's float_to_half_full (3rdParty/tinyexr/tinyexr.h:7044-7080, after
// the ISPC reference code): float denormals -> signed zero; Inf / NaN -> Inf
// / quiet NaN; overflow -> Inf; otherwise the mantissa is truncated and
// rounded up when the first dropped bit is set (ties away from zero), the
// carry running into the exponent; half denormals likewise.
void fornos_float_to_half(const float* in, uint16_t* out, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        uint32_t f;
        std::memcpy(&f, &in[i], 4);
        const uint32_t sign = (f >> 16) & 0x8000u, e = (f >> 23) & 0xffu, m = f & 0x7fffffu;
        uint32_t h = 0;
        if (e == 0) {
            h = 0;
        } else if (e == 255) {
            h = (31u << 10) | (m ? 0x200u : 0u);
        } else {
            const int ne = (int)e - 127 + 15;
            if (ne >= 31) {
                h = 31u << 10;
            } else if (ne <= 0) {
                if (14 - ne <= 24) {
                    const uint32_t mant = m | 0x800000u;
                    h = mant >> (14 - ne);
                    if ((mant >> (13 - ne)) & 1u) h += 1;
                }
            } else {
                h = ((uint32_t)ne << 10) | (m >> 13);
                if (m & 0x1000u) h += 1;
            }
        }
        out[i] = (uint16_t)((h & 0x7fffu) | sign);
    }
}

fornos_status fornos_write_png(const char* path, uint32_t width, uint32_t height, const uint8_t* rgb)
{
    if (!path || !rgb || width == 0 || height == 0) return set_error(FORNOS_INVALID_ARG, "fornos_write_png: bad argument");
    std::vector<uint8_t> z;
    if (!png_idat(rgb, width, height, 6, z)) return set_error(FORNOS_OOM, "fornos_write_png: deflate failed");
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    be32(ihdr, width);
    be32(ihdr, height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8 bits, RGB, deflate, adaptive filters, no interlace
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", {});
    File f(path);
    if (!f.f || !f.put(out.data(), out.size())) return set_error(FORNOS_INVALID_ARG, std::string("cannot write ") + path);
    return FORNOS_OK;
}

// Uncompressed true-colour TGA, first row at the top (descriptor bit 5).
fornos_status fornos_write_tga(const char* path, uint32_t width, uint32_t height, const uint8_t* rgb)
{
    if (!path || !rgb || width == 0 || height == 0 || width > 65535 || height > 65535)
        return set_error(FORNOS_INVALID_ARG, "fornos_write_tga: bad argument");
    std::vector<uint8_t> out(18, 0);
    out[2] = 2;
    out[12] = (uint8_t)width;
    out[13] = (uint8_t)(width >> 8);
    out[14] = (uint8_t)height;
    out[15] = (uint8_t)(height >> 8);
    out[16] = 24;
    out[17] = 0x20;
    out.reserve(18 + (size_t)width * height * 3);
    for (size_t i = 0; i < (size_t)width * height; ++i) {
        out.push_back(rgb[3 * i + 2]);
        out.push_back(rgb[3 * i + 1]);
        out.push_back(rgb[3 * i + 0]);
    }
    File f(path);
    if (!f.f || !f.put(out.data(), out.size())) return set_error(FORNOS_INVALID_ARG, std::string("cannot write ") + path);
    return FORNOS_OK;
}

// Scanline OpenEXR, no compression, increasing y, half channels in the given
// order (names must be sorted, as the format requires).
fornos_status fornos_write_exr(const char* path, uint32_t width, uint32_t height, uint32_t channels,
                               const char* const* names, const float* const* planes)
{
    if (!path || !names || !planes || width == 0 || height == 0 || channels == 0 || channels > 16)
        return set_error(FORNOS_INVALID_ARG, "fornos_write_exr: bad argument");
    for (uint32_t c = 1; c < channels; ++c)
        if (std::strcmp(names[c - 1], names[c]) >= 0) return set_error(FORNOS_INVALID_ARG, "fornos_write_exr: channel names must be sorted");
    std::vector<uint8_t> h = {0x76, 0x2f, 0x31, 0x01, 2, 0, 0, 0};
    std::vector<uint8_t> cl;
    for (uint32_t c = 0; c < channels; ++c) {
        cl.insert(cl.end(), names[c], names[c] + std::strlen(names[c]) + 1);
        le(cl, 1, 4);        // HALF
        le(cl, 0, 4);        // pLinear + reserved
        le(cl, 1, 4);        // xSampling
        le(cl, 1, 4);        // ySampling
    }
    cl.push_back(0);
    exr_attr(h, "channels", "chlist", cl);
    exr_attr(h, "compression", "compression", {0});
    std::vector<uint8_t> box;
    le(box, 0, 4);
    le(box, 0, 4);
    le(box, width - 1, 4);
    le(box, height - 1, 4);
    exr_attr(h, "dataWindow", "box2i", box);
    exr_attr(h, "displayWindow", "box2i", box);
    exr_attr(h, "lineOrder", "lineOrder", {0});
    const float one = 1.0f, zero = 0.0f;
    std::vector<uint8_t> f1(4), v2(8);
    std::memcpy(f1.data(), &one, 4);
    std::memcpy(v2.data(), &zero, 4);
    std::memcpy(v2.data() + 4, &zero, 4);
    exr_attr(h, "pixelAspectRatio", "float", f1);
    exr_attr(h, "screenWindowCenter", "v2f", v2);
    exr_attr(h, "screenWindowWidth", "float", f1);
    h.push_back(0);
    const size_t line = (size_t)channels * width * 2, block = 8 + line;
    const size_t table_at = h.size(), data_at = table_at + 8 * (size_t)height;
    for (uint32_t y = 0; y < height; ++y) le(h, data_at + (size_t)y * block, 8);
    std::vector<uint16_t> half(width);
    h.reserve(data_at + (size_t)height * block);
    for (uint32_t y = 0; y < height; ++y) {
        le(h, y, 4);
        le(h, line, 4);
        for (uint32_t c = 0; c < channels; ++c) {
            fornos_float_to_half(planes[c] + (size_t)y * width, half.data(), width);
            for (uint32_t x = 0; x < width; ++x) le(h, half[x], 2);
        }
    }
    File f(path);
    if (!f.f || !f.put(h.data(), h.size())) return set_error(FORNOS_INVALID_ARG, std::string("cannot write ") + path);
    return FORNOS_OK;
}

fornos_status fornos_export_float_image(const fornos_solver* s, const char* path, float filter_lo, float filter_hi,
                                        int normalize, int dilate, float* minmax)
{
    if (!s || !path) return set_error(FORNOS_INVALID_ARG, "fornos_export_float_image: null argument");
    uint32_t w = 0, h = 0;
    if (fornos_status e = fornos_solver_map_size(s, &w, &h)) return e;
    switch (extension(path)) {
    case Ext::Png:
    case Ext::Tga: {
        std::vector<uint8_t> rgb((size_t)w * h * 3);
        if (fornos_status e = fornos_export_float_rgb8(s, filter_lo, filter_hi, normalize, dilate, rgb.data(), minmax))
            return e;
        return extension(path) == Ext::Png ? fornos_write_png(path, w, h, rgb.data()) : fornos_write_tga(path, w, h, rgb.data());
    }
    case Ext::Exr: {
        std::vector<float> img((size_t)w * h);
        if (fornos_status e = fornos_export_float_f32(s, filter_lo, filter_hi, normalize, img.data(), minmax)) return e;
        const char* names[] = {"B"};
        const float* planes[] = {img.data()};
        return fornos_write_exr(path, w, h, 1, names, planes);
    }
    default:
        return set_error(FORNOS_INVALID_ARG, std::string("unsupported image extension: ") + path);
    }
}

fornos_status fornos_export_vector_image(const fornos_solver* s, const char* path)
{
    if (!s || !path) return set_error(FORNOS_INVALID_ARG, "fornos_export_vector_image: null argument");
    if (extension(path) != Ext::Exr) return set_error(FORNOS_INVALID_ARG, "fornos_export_vector_image: EXR only");
    uint32_t w = 0, h = 0;
    if (fornos_status e = fornos_solver_map_size(s, &w, &h)) return e;
    std::vector<float> planes((size_t)w * h * 3);
    if (fornos_status e = fornos_export_vector_f32(s, planes.data())) return e;
    const char* names[] = {"B", "G", "R"};
    const float* p[] = {planes.data(), planes.data() + (size_t)w * h, planes.data() + 2 * (size_t)w * h};
    return fornos_write_exr(path, w, h, 3, names, p);
}

fornos_status fornos_export_normal_image(const fornos_solver* s, const char* path, int dilate)
{
    if (!s || !path) return set_error(FORNOS_INVALID_ARG, "fornos_export_normal_image: null argument");
    const Ext e = extension(path);
    if (e == Ext::Exr) return fornos_export_vector_image(s, path);
    if (e == Ext::Unknown) return set_error(FORNOS_INVALID_ARG, std::string("unsupported image extension: ") + path);
    uint32_t w = 0, h = 0;
    if (fornos_status st = fornos_solver_map_size(s, &w, &h)) return st;
    std::vector<uint8_t> rgb((size_t)w * h * 3);
    if (fornos_status st = fornos_export_normal_rgb8(s, dilate, rgb.data())) return st;
    return e == Ext::Png ? fornos_write_png(path, w, h, rgb.data()) : fornos_write_tga(path, w, h, rgb.data());
}

}  // extern "C"
