// orf_vec.h -- small FP32 vector types shared by host C++ and gfx950 kernels.
//
// Conventions follow the OptiX math the reference programs use: normalize()
// multiplies by the reciprocal length, faceforward(n, i, nref) = n * sign(i.nref).
// All device code in this library is compiled with -ffp-contract=off so that
// every +,-,*,/ and sqrt rounds exactly as the host restatement does; the
// expression order of each formula is part of its parity contract.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "orf_math.h"

#define ORF_HD __host__ __device__ __forceinline__

namespace orf {

struct f2 { float x, y; };
struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

ORF_HD f2 mk2(float x, float y) { return f2{x, y}; }
ORF_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
ORF_HD f3 mk3(float s) { return f3{s, s, s}; }

ORF_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
ORF_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
ORF_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
ORF_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
ORF_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
ORF_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
ORF_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
ORF_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
ORF_HD f3 operator+(f3 a, float s) { return f3{a.x + s, a.y + s, a.z + s}; }
ORF_HD f3 operator+(float s, f3 a) { return f3{s + a.x, s + a.y, s + a.z}; }
ORF_HD f3 operator-(f3 a, float s) { return f3{a.x - s, a.y - s, a.z - s}; }
ORF_HD f3 operator-(float s, f3 a) { return f3{s - a.x, s - a.y, s - a.z}; }
ORF_HD f3& operator+=(f3& a, f3 b) { a = a + b; return a; }
ORF_HD f3& operator*=(f3& a, f3 b) { a = a * b; return a; }
ORF_HD f3& operator*=(f3& a, float s) { a = a * s; return a; }

ORF_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
ORF_HD f3 cross(f3 a, f3 b)
{
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
ORF_HD f3 normalize(f3 v)
{
    float inv = 1.0f / sqrtf(dot(v, v));
    return v * inv;
}
ORF_HD float length(f3 v) { return sqrtf(dot(v, v)); }
ORF_HD f3 faceforward(f3 n, f3 i, f3 nref) { return n * copysignf(1.0f, dot(i, nref)); }
ORF_HD f3 vmin(f3 a, f3 b) { return f3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
ORF_HD f3 vmax(f3 a, f3 b) { return f3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
ORF_HD bool finite3(f3 v) { return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z); }
ORF_HD uint32_t f2u(float f)
{
    union {
        float f;
        uint32_t u;
    } c;
    c.f = f;
    return c.u;
}
ORF_HD float u2f(uint32_t u) { return orf_u2f(u); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;

}  // namespace orf
