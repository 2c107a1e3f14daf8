// orf_assets.h -- host-side asset readers of the scene loader: a JSON document
// model, Wavefront OBJ meshes and Radiance RGBE (.hdr) images (internal).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace orf {

// ------------------------------------------------------------------ JSON

struct JsonValue {
    enum Kind { NUL, BOOL, NUMBER, STRING, ARRAY, OBJECT } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JsonValue> arr;
    std::map<std::string, JsonValue> obj;

    bool is_object() const { return kind == OBJECT; }
    bool is_array() const { return kind == ARRAY; }
    bool is_number() const { return kind == NUMBER; }
    bool is_string() const { return kind == STRING; }
    // Member `key` of an object, or nullptr.
    const JsonValue* get(const std::string& key) const;
};

// Parses RFC 8259 JSON text; throws std::runtime_error with the byte offset on
// malformed input.
JsonValue json_parse(const std::string& text);

// JSON text of v (objects with sorted keys, two-space indentation, numbers as
// %.17g so every double reads back exactly).
std::string json_write(const JsonValue& v);

// ------------------------------------------------------------------ OBJ

// Triangles of every face of the file (polygons as fans v0, vi, vi+1), indices
// 0-based into verts / normals.  nidx is empty when some face has no normals.
struct ObjMesh {
    std::vector<float> verts;    // 3 per vertex
    std::vector<float> normals;  // 3 per normal
    std::vector<int32_t> tris;   // 3 per triangle
    std::vector<int32_t> nidx;   // 3 per triangle, or empty
};

ObjMesh obj_read(const std::string& path);

// ------------------------------------------------------------------ HDR

// RGBA FP32, row 0 = the bottom row of the file (the file stores -Y +X, top
// first), alpha 1.
struct HdrImage {
    int width = 0, height = 0;
    std::vector<float> rgba;
};

HdrImage hdr_read(const std::string& path);

}  // namespace orf
