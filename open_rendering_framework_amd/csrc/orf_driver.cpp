// orf_driver.cpp -- progressive frame driver and film files on the C-ABI (host).
//
// Restates OptixScene's frame loop and film I/O: renderScene (one launch per
// frame after the translucent probe pass, ORF/OptixScene.cpp:139-168),
// saveScene (scene document + PNG + .raw, :202-236; the document through
// OptixSceneLoader::saveJSONScene, ORF/OptixSceneLoader.cpp:48-70) and
// loadBuffer (.raw into the film, ORF/OptixScene.cpp:251-299).
//
// The film here accumulates FP64 sums and sample counts; its mean is what the
// reference's FP32 running-mean output_buffer holds (CU/path_tracer.cu:64-68),
// and .raw files are written from and read into that mean.  A .raw is
// width x height FP32 RGBA, row 0 = launch row 0 (the bottom of the image),
// 16 bytes per pixel, no header.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "orf_assets.h"
#include "orf_common.h"
#include "orf_context.h"

namespace orf {
namespace {

// Mean RGBA (alpha 1) of the film, FP32, launch-row order.
std::vector<float> film_mean(orf_context_t* c)
{
    ORF_REQUIRE(c->d_film.p && c->film_w > 0, "film: nothing rendered yet");
    const size_t n = (size_t)c->film_w * c->film_h;
    std::vector<double> s(4 * n);
    ORF_HIP_CHECK(hipDeviceSynchronize());
    ORF_HIP_CHECK(hipMemcpy(s.data(), c->d_film.p, s.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double> dep;
    if (c->integrator == 1) {
        dep.resize(n);
        ORF_HIP_CHECK(hipMemcpy(dep.data(), c->d_depth.p, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    std::vector<float> m(4 * n);
    for (size_t i = 0; i < n; ++i) {
        const double cnt = s[4 * i + 3];
        for (int k = 0; k < 3; ++k) m[4 * i + k] = (float)(cnt > 0 ? s[4 * i + k] / cnt : 0.0);
        m[4 * i + 3] = dep.empty() ? 1.0f : (float)(cnt > 0 ? dep[i] / cnt : 0.0);  // depth tracer: hit distance
    }
    return m;
}

void write_file(const std::string& path, const void* data, size_t bytes)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    ORF_REQUIRE(f, "cannot write '" + path + "'");
    const size_t w = std::fwrite(data, 1, bytes, f);
    const int rc = std::fclose(f);
    ORF_REQUIRE(w == bytes && rc == 0, "short write to '" + path + "'");
}

uint32_t crc32(const uint8_t* p, size_t n, uint32_t crc = 0)
{
    // CRC-32 of ISO 3309 / ITU-T V.42 (the PNG chunk checksum), bitwise
    crc = ~crc;
    for (size_t i = 0; i < n; ++i) {
        crc ^= p[i];
        for (int k = 0; k < 8; ++k) crc = (crc >> 1) ^ (0xEDB88320u & (0u - (crc & 1u)));
    }
    return ~crc;
}

void put_be32(std::vector<uint8_t>& out, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) out.push_back((uint8_t)(v >> s));
}

void png_chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data)
{
    put_be32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put_be32(out, crc32(out.data() + start, out.size() - start));
}

// 8-bit RGBA PNG (RFC 2083) with uncompressed deflate blocks (RFC 1951 type 0)
// in a zlib stream (RFC 1950).  rows: top row first.
std::vector<uint8_t> png_encode(const std::vector<uint8_t>& rgba, int w, int h)
{
    std::vector<uint8_t> raw;
    raw.reserve((size_t)h * (4 * (size_t)w + 1));
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);  // filter: none
        raw.insert(raw.end(), rgba.begin() + (size_t)y * 4 * w, rgba.begin() + (size_t)(y + 1) * 4 * w);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    for (size_t off = 0; off < raw.size() || off == 0;) {
        const size_t len = std::min<size_t>(65535, raw.size() - off);
        const bool last = off + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)(len & 0xFF));
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)(~len & 0xFF));
        z.push_back((uint8_t)((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
        off += len;
        if (last) break;
    }
    uint32_t a = 1, b = 0;  // Adler-32
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)w);
    put_be32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8 bits, RGBA, deflate, no filter choice, no interlace
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", {});
    return out;
}

// bufferToImage for FLOAT4 (ORF/ImageLoader.cpp:110-133): rows flipped to top
// first, each channel (int)(v * 255) clamped to [0, 255].
std::vector<uint8_t> to_rgba8(const std::vector<float>& m, int w, int h)
{
    std::vector<uint8_t> out((size_t)w * h * 4);
    for (int j = h - 1; j >= 0; --j)
        for (int i = 0; i < w; ++i)
            for (int k = 0; k < 4; ++k) {
                const float v = m[4 * ((size_t)j * w + i) + k] * 255.0f;
                int P = std::isnan(v) ? 0 : (v >= 2147483520.0f ? 255 : (v <= -2147483520.0f ? 0 : (int)v));
                out[4 * ((size_t)(h - 1 - j) * w + i) + k] = (uint8_t)(P < 0 ? 0 : P > 255 ? 255 : P);
            }
    return out;
}

std::string strip_extension(const std::string& path)
{
    const size_t slash = path.find_last_of('/');
    const size_t dot = path.find_last_of('.');
    return dot != std::string::npos && (slash == std::string::npos || dot > slash) ? path.substr(0, dot) : path;
}

}  // namespace
}  // namespace orf

using namespace orf;

extern "C" orf_status orf_render_frames(orf_context ctx, int width, int height, uint32_t frame_begin,
                                        uint32_t frame_end, void* stream)
{
    ORF_API_BEGIN
    ORF_REQUIRE(frame_end >= frame_begin, "orf_render_frames: frame_end before frame_begin");
    for (uint32_t f = frame_begin; f < frame_end; ++f) {
        const orf_status s = orf_launch(ctx, width, height, f, 1, stream);
        if (s != ORF_SUCCESS) return s;
    }
    ORF_API_END
}

extern "C" orf_status orf_film_save_raw(orf_context ctx, const char* path)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(path, "orf_film_save_raw: null path");
    const std::vector<float> m = film_mean(c);
    write_file(path, m.data(), m.size() * sizeof(float));
    ORF_API_END
}

extern "C" orf_status orf_film_load_raw(orf_context ctx, const char* path, int width, int height, uint32_t frames)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(path && width > 0 && height > 0 && width <= 16384 && height <= 16384 && frames > 0,
                "orf_film_load_raw: bad arguments");
    const size_t n = (size_t)width * height;
    std::vector<float> m(4 * n);
    FILE* f = std::fopen(path, "rb");
    if (!f) ORF_THROW(ORF_ERROR_NOT_FOUND, std::string("cannot open '") + path + "'");
    const size_t got = std::fread(m.data(), sizeof(float), m.size(), f);
    const bool extra = std::fgetc(f) != EOF;
    std::fclose(f);
    if (got != m.size() || extra)
        ORF_THROW(ORF_ERROR_FILE_FORMAT, std::string("'") + path + "' is not a " + std::to_string(width) + " x " +
                                             std::to_string(height) + " FP32 RGBA film");
    std::vector<double> s(4 * n), dep(n);
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) s[4 * i + k] = (double)m[4 * i + k] * (double)frames;
        s[4 * i + 3] = (double)frames;
        dep[i] = (double)m[4 * i + 3] * (double)frames;  // the depth tracer's alpha
    }
    film_ensure(c, width, height);
    ORF_HIP_CHECK(hipDeviceSynchronize());
    ORF_HIP_CHECK(hipMemcpy(c->d_film.p, s.data(), s.size() * sizeof(double), hipMemcpyHostToDevice));
    if (c->integrator == 1)
        ORF_HIP_CHECK(hipMemcpy(c->d_depth.p, dep.data(), dep.size() * sizeof(double), hipMemcpyHostToDevice));
    ORF_API_END
}

extern "C" orf_status orf_film_save_png(orf_context ctx, const char* path)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(path, "orf_film_save_png: null path");
    const std::vector<uint8_t> png = png_encode(to_rgba8(film_mean(c), c->film_w, c->film_h), c->film_w, c->film_h);
    write_file(path, png.data(), png.size());
    ORF_API_END
}

extern "C" orf_status orf_scene_save(orf_context ctx, const char* path, uint32_t frame_count)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(path, "orf_scene_save: null path");
    ORF_REQUIRE(!c->scene_doc.empty(), "orf_scene_save: no scene document was loaded");
    const std::string base = strip_extension(path);
    const size_t slash = base.find_last_of('/');
    JsonValue doc = json_parse(c->scene_doc);
    JsonValue fc, bp;
    fc.kind = JsonValue::NUMBER;
    fc.num = (double)frame_count;
    bp.kind = JsonValue::STRING;
    bp.str = (slash == std::string::npos ? base : base.substr(slash + 1)) + ".raw";  // beside the document
    doc.obj["frame_count"] = fc;
    doc.obj["buffer_path"] = bp;
    const std::string text = json_write(doc);
    write_file(base + ".json", text.data(), text.size());
    const std::vector<float> m = film_mean(c);
    const std::vector<uint8_t> png = png_encode(to_rgba8(m, c->film_w, c->film_h), c->film_w, c->film_h);
    write_file(base + ".png", png.data(), png.size());
    write_file(base + ".raw", m.data(), m.size() * sizeof(float));
    ORF_API_END
}
