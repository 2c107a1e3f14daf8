// orf_assets.cpp -- JSON, Wavefront OBJ and Radiance RGBE readers (host).
//
// They replace the reference's Qt JSON document (QJsonDocument, used by
// ORF/OptixSceneLoader.cpp:26-45), the GLM OBJ parser with
// ObjLoader::loadVertexData / createGeometryInstances (ORF/glm.cpp,
// ORF/ObjLoader.cpp:198-330) and HDRLoader (ORF/HDRLoader.cpp:63-181, with the
// row flip of loadHDRTexture, :263-281), written from the file formats'
// published descriptions.
#include "orf_assets.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "orf_common.h"

namespace orf {

// ------------------------------------------------------------------ JSON

const JsonValue* JsonValue::get(const std::string& key) const
{
    if (kind != OBJECT) return nullptr;
    auto it = obj.find(key);
    return it == obj.end() ? nullptr : &it->second;
}

namespace {

struct JsonParser {
    const std::string& s;
    size_t i = 0;

    [[noreturn]] void fail(const char* what) const
    {
        throw std::runtime_error(std::string("json: ") + what + " at byte " + std::to_string(i));
    }
    void ws()
    {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool lit(const char* w)
    {
        size_t n = std::strlen(w);
        if (s.compare(i, n, w) != 0) return false;
        i += n;
        return true;
    }
    static void put_utf8(std::string& out, unsigned cp)
    {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    unsigned hex4()
    {
        if (i + 4 > s.size()) fail("short \\u escape");
        unsigned v = 0;
        for (int k = 0; k < 4; ++k) {
            char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string string()
    {
        if (s[i] != '"') fail("expected string");
        ++i;
        std::string out;
        for (;;) {
            if (i >= s.size()) fail("unterminated string");
            char c = s[i++];
            if (c == '"') break;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= s.size()) fail("unterminated escape");
            char e = s[i++];
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                unsigned cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00 && i + 6 <= s.size() && s[i] == '\\' && s[i + 1] == 'u') {
                    i += 2;
                    unsigned lo = hex4();
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                put_utf8(out, cp);
                break;
            }
            default: fail("bad escape");
            }
        }
        return out;
    }
    double number()
    {
        size_t st = i;
        if (s[i] == '-') ++i;
        if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9')) fail("bad number");
        while (i < s.size() && ((s[i] >= '0' && s[i] <= '9') || s[i] == '.' || s[i] == 'e' || s[i] == 'E' ||
                                s[i] == '+' || s[i] == '-'))
            ++i;
        std::string tok = s.substr(st, i - st);
        char* end = nullptr;
        double v = std::strtod(tok.c_str(), &end);
        if (!end || *end != '\0') fail("bad number");
        return v;
    }
    JsonValue value(int depth)
    {
        if (depth > 256) fail("nesting too deep");
        ws();
        if (i >= s.size()) fail("unexpected end");
        JsonValue v;
        char c = s[i];
        if (c == '{') {
            v.kind = JsonValue::OBJECT;
            ++i;
            ws();
            if (i < s.size() && s[i] == '}') {
                ++i;
                return v;
            }
            for (;;) {
                ws();
                if (i >= s.size()) fail("unterminated object");
                std::string k = string();
                ws();
                if (i >= s.size() || s[i] != ':') fail("expected ':'");
                ++i;
                v.obj[k] = value(depth + 1);
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    ++i;
                    break;
                }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = JsonValue::ARRAY;
            ++i;
            ws();
            if (i < s.size() && s[i] == ']') {
                ++i;
                return v;
            }
            for (;;) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == ']') {
                    ++i;
                    break;
                }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = JsonValue::STRING;
            v.str = string();
        } else if (lit("true")) {
            v.kind = JsonValue::BOOL;
            v.b = true;
        } else if (lit("false")) {
            v.kind = JsonValue::BOOL;
        } else if (lit("null")) {
            v.kind = JsonValue::NUL;
        } else {
            v.kind = JsonValue::NUMBER;
            v.num = number();
        }
        return v;
    }
};

std::string read_file(const std::string& path, bool binary)
{
    std::ifstream f(path, binary ? std::ios::binary : std::ios::in);
    ORF_REQUIRE(f.is_open(), "cannot open '" + path + "'");
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

}  // namespace

JsonValue json_parse(const std::string& text)
{
    JsonParser p{text};
    JsonValue v = p.value(0);
    p.ws();
    if (p.i != text.size()) p.fail("trailing characters");
    return v;
}

namespace {

void json_string(std::string& out, const std::string& s)
{
    out += '"';
    for (unsigned char ch : s) {
        switch (ch) {
        case '"': out += "\\\""; break;
        case '\\': out += "\\\\"; break;
        case '\b': out += "\\b"; break;
        case '\f': out += "\\f"; break;
        case '\n': out += "\\n"; break;
        case '\r': out += "\\r"; break;
        case '\t': out += "\\t"; break;
        default:
            if (ch < 0x20) {
                char buf[8];
                std::snprintf(buf, sizeof(buf), "\\u%04x", ch);
                out += buf;
            } else {
                out += (char)ch;
            }
        }
    }
    out += '"';
}

void json_emit(std::string& out, const JsonValue& v, int depth)
{
    const std::string pad(2 * (size_t)(depth + 1), ' '), end(2 * (size_t)depth, ' ');
    switch (v.kind) {
    case JsonValue::NUL: out += "null"; break;
    case JsonValue::BOOL: out += v.b ? "true" : "false"; break;
    case JsonValue::NUMBER: {
        char buf[32];
        std::snprintf(buf, sizeof(buf), "%.17g", v.num);  // round-trips every double
        out += buf;
        break;
    }
    case JsonValue::STRING: json_string(out, v.str); break;
    case JsonValue::ARRAY: {
        out += '[';
        for (size_t i = 0; i < v.arr.size(); ++i) {
            if (i) out += ", ";
            json_emit(out, v.arr[i], depth + 1);
        }
        out += ']';
        break;
    }
    case JsonValue::OBJECT: {
        out += "{\n";
        size_t i = 0;
        for (const auto& kv : v.obj) {
            out += pad;
            json_string(out, kv.first);
            out += ": ";
            json_emit(out, kv.second, depth + 1);
            out += ++i < v.obj.size() ? ",\n" : "\n";
        }
        out += end + "}";
        break;
    }
    }
}

}  // namespace

std::string json_write(const JsonValue& v)
{
    std::string out;
    json_emit(out, v, 0);
    out += '\n';
    return out;
}

// ------------------------------------------------------------------ OBJ

namespace {

// 1-based or negative (relative) OBJ index -> 0-based, checked against n.
int obj_index(long v, size_t n, const std::string& path, int line)
{
    long k = v > 0 ? v - 1 : (long)n + v;
    if (v == 0 || k < 0 || k >= (long)n)
        throw std::runtime_error("obj '" + path + "' line " + std::to_string(line) + ": index out of range");
    return (int)k;
}

}  // namespace

ObjMesh obj_read(const std::string& path)
{
    std::string text = read_file(path, false);
    ObjMesh m;
    std::vector<int32_t> nidx;
    bool all_normals = true;
    std::istringstream in(text);
    std::string line;
    int ln = 0;
    while (std::getline(in, line)) {
        ++ln;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        size_t p = line.find_first_not_of(" \t");
        if (p == std::string::npos || line[p] == '#') continue;
        std::istringstream ls(line.substr(p));
        std::string tag;
        ls >> tag;
        if (tag == "v") {
            float x = 0, y = 0, z = 0;
            ls >> x >> y >> z;
            m.verts.insert(m.verts.end(), {x, y, z});
        } else if (tag == "vn") {
            float x = 0, y = 0, z = 0;
            ls >> x >> y >> z;
            m.normals.insert(m.normals.end(), {x, y, z});
        } else if (tag == "f") {
            std::vector<int> fv, fn;
            std::string tok;
            bool face_normals = true;
            while (ls >> tok) {
                // v, v/t, v//n or v/t/n
                long vi = 0, ni = 0;
                size_t s1 = tok.find('/');
                vi = std::strtol(tok.substr(0, s1).c_str(), nullptr, 10);
                if (s1 != std::string::npos) {
                    size_t s2 = tok.find('/', s1 + 1);
                    if (s2 != std::string::npos && s2 + 1 < tok.size())
                        ni = std::strtol(tok.substr(s2 + 1).c_str(), nullptr, 10);
                }
                fv.push_back(obj_index(vi, m.verts.size() / 3, path, ln));
                if (ni != 0) fn.push_back(obj_index(ni, m.normals.size() / 3, path, ln));
                else face_normals = false;
            }
            if (fv.size() < 3)
                throw std::runtime_error("obj '" + path + "' line " + std::to_string(ln) + ": face with < 3 vertices");
            all_normals = all_normals && face_normals;
            for (size_t k = 1; k + 1 < fv.size(); ++k) {
                m.tris.insert(m.tris.end(), {fv[0], fv[k], fv[k + 1]});
                if (face_normals) nidx.insert(nidx.end(), {fn[0], fn[k], fn[k + 1]});
                else nidx.insert(nidx.end(), {0, 0, 0});
            }
        }
        // vt, g, o, s, usemtl, mtllib: not used by the render path
    }
    if (all_normals && !m.normals.empty()) m.nidx = std::move(nidx);
    return m;
}

// ------------------------------------------------------------------ HDR

namespace {

struct ByteReader {
    const std::string& d;
    size_t i = 0;
    int get()
    {
        if (i >= d.size()) throw std::runtime_error("hdr: premature end of file");
        return (unsigned char)d[i++];
    }
};

}  // namespace

HdrImage hdr_read(const std::string& path)
{
    std::string data = read_file(path, true);
    ByteReader r{data};
    auto line = [&]() {
        std::string l;
        for (;;) {
            int c = r.get();
            if (c == '\n') break;
            l += (char)c;
        }
        if (!l.empty() && l.back() == '\r') l.pop_back();
        return l;
    };
    std::string magic = line();
    if (magic != "#?RADIANCE" && magic != "#?RGBE")
        throw std::runtime_error("hdr '" + path + "': not a Radiance file");
    float exposure = 1.0f;
    for (;;) {
        std::string h = line();
        if (h.empty()) break;
        if (h[0] == '#') continue;
        // a FORMAT line must name RGBE; the last line holding EXPOSURE= sets the
        // exposure (the header loop of ORF/HDRLoader.cpp:143-161, which does not
        // multiply repeated EXPOSURE lines as the Radiance format would)
        if (h.find("FORMAT") != std::string::npos) {
            if (h != "FORMAT=32-bit_rle_rgbe") throw std::runtime_error("hdr '" + path + "': only RGBE is read");
            continue;
        }
        const size_t ofs = h.find("EXPOSURE=");
        if (ofs != std::string::npos) exposure = (float)std::atof(h.c_str() + ofs + 9);
    }
    std::string res = line();
    char ya[3] = {0}, xa[3] = {0};
    int ny = 0, nx = 0;
    if (std::sscanf(res.c_str(), "%2s %d %2s %d", ya, &ny, xa, &nx) != 4 || std::string(ya) != "-Y" ||
        std::string(xa) != "+X" || nx <= 0 || ny <= 0)
        throw std::runtime_error("hdr '" + path + "': only -Y N +X M images are read");
    std::vector<unsigned char> rgbe((size_t)nx * ny * 4);
    for (int y = 0; y < ny; ++y) {
        unsigned char* row = rgbe.data() + (size_t)y * nx * 4;
        size_t save = r.i;
        bool rle = false;
        if (nx >= 8 && nx <= 0x7fff) {
            int c0 = r.get(), c1 = r.get(), c2 = r.get(), c3 = r.get();
            if (c0 == 2 && c1 == 2 && !(c2 & 0x80)) {
                if (((c2 << 8) | c3) != nx) throw std::runtime_error("hdr '" + path + "': scanline width mismatch");
                rle = true;
            } else {
                r.i = save;
            }
        }
        if (!rle) {  // flat scanline (4 bytes per pixel)
            for (int x = 0; x < nx * 4; ++x) row[x] = (unsigned char)r.get();
            continue;
        }
        for (int ch = 0; ch < 4; ++ch)
            for (int x = 0; x < nx;) {
                int code = r.get();
                if (code > 128) {  // run
                    int n = code & 127, v = r.get();
                    if (x + n > nx) throw std::runtime_error("hdr '" + path + "': run past scanline end");
                    while (n--) row[4 * (x++) + ch] = (unsigned char)v;
                } else {  // literal span
                    if (code == 0 || x + code > nx) throw std::runtime_error("hdr '" + path + "': bad span");
                    while (code--) row[4 * (x++) + ch] = (unsigned char)r.get();
                }
            }
    }
    HdrImage img;
    img.width = nx;
    img.height = ny;
    img.rgba.resize((size_t)nx * ny * 4);
    const float inv_exposure = 1.0f / exposure;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const unsigned char* p = rgbe.data() + ((size_t)(ny - 1 - y) * nx + x) * 4;
            float* o = img.rgba.data() + ((size_t)y * nx + x) * 4;
            if (p[3] == 0) {
                o[0] = o[1] = o[2] = 0.0f;
            } else {
                // value = (mantissa + 1/2) 2^(e - 136), the Radiance convention
                float s = (float)std::ldexp(1.0, (int)p[3] - (128 + 8)) * inv_exposure;
                o[0] = (p[0] + 0.5f) * s;
                o[1] = (p[1] + 0.5f) * s;
                o[2] = (p[2] + 0.5f) * s;
            }
            o[3] = 1.0f;
        }
    return img;
}

}  // namespace orf
