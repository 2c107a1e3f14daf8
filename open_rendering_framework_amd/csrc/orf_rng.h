// orf_rng.h -- per-path random numbers (device and host).
//
// The render path keeps one 32-bit state per path and advances it with 16
// TEA rounds per draw, exactly the stream of ORF/random.h:25-40 (tea<N>) and
// :117-126 (tea_hash / rnd_tea); the pixel seed is tea<16>(W*y + x, frame)
// (CU/path_tracer.cu:48).  Bit-exactness of this stream is what lets the
// oracle and the kernels follow the same paths.
//
// The translucent programs draw from a second, 64-bit stream (ORF/random.h:128-142,
// rnd_accurate, under RND_64): the state (x, y) is replaced by the first two
// words of the MD5 digest (RFC 1321) of its own 8 bytes, and the draw is the
// digest's low 63 bits over 2^63, rounded to double and then to float.  The
// camera seeds both words with the pixel seed (CU/path_tracer.cu:49), the probe
// program with the probe seed (CU/sample_camera.cu:55-57).
#pragma once

#include "orf_vec.h"

namespace orf {

template <unsigned N>
ORF_HD uint32_t tea(uint32_t v0, uint32_t v1)
{
    uint32_t s0 = 0;
#pragma unroll
    for (unsigned n = 0; n < N; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

// Uniform in [0,1] (float rounding can give exactly 1.0f, as in the reference).
ORF_HD float rnd_tea(uint32_t& state)
{
    state = tea<16>(state, 100u);
    return (float)(state & 0x7FFFFFFFu) / (float)0x80000000u;
}

// Numerical Recipes LCG with 24-bit output (ORF/random.h:43-61).
ORF_HD float rnd_lcg(uint32_t& state)
{
    state = 1664525u * state + 1013904223u;
    return (float)(state & 0x00FFFFFFu) / (float)0x01000000u;
}

// 64-bit random state of the translucent programs (Seed64 of ORF/random.h:128-132).
struct Seed64 {
    uint32_t x, y;
};

ORF_HD Seed64 seed64(uint32_t s) { return Seed64{s, s}; }

ORF_HD uint32_t rotl32(uint32_t v, int n) { return (v << n) | (v >> (32 - n)); }

// MD5 (RFC 1321) of the 8-byte message holding x then y (little-endian words).
// The padded message is one 512-bit block: words x, y, 0x80, zeros, the bit
// length 64 in word 14 and 0 in word 15.  Returns the digest words a, b, c, d.
ORF_HD uint4 md5_8(uint32_t x, uint32_t y)
{
    // K[i] = floor(|sin(i + 1)| * 2^32)
    constexpr uint32_t K[64] = {
        0xd76aa478u, 0xe8c7b756u, 0x242070dbu, 0xc1bdceeeu, 0xf57c0fafu, 0x4787c62au, 0xa8304613u, 0xfd469501u,
        0x698098d8u, 0x8b44f7afu, 0xffff5bb1u, 0x895cd7beu, 0x6b901122u, 0xfd987193u, 0xa679438eu, 0x49b40821u,
        0xf61e2562u, 0xc040b340u, 0x265e5a51u, 0xe9b6c7aau, 0xd62f105du, 0x02441453u, 0xd8a1e681u, 0xe7d3fbc8u,
        0x21e1cde6u, 0xc33707d6u, 0xf4d50d87u, 0x455a14edu, 0xa9e3e905u, 0xfcefa3f8u, 0x676f02d9u, 0x8d2a4c8au,
        0xfffa3942u, 0x8771f681u, 0x6d9d6122u, 0xfde5380cu, 0xa4beea44u, 0x4bdecfa9u, 0xf6bb4b60u, 0xbebfbc70u,
        0x289b7ec6u, 0xeaa127fau, 0xd4ef3085u, 0x04881d05u, 0xd9d4d039u, 0xe6db99e5u, 0x1fa27cf8u, 0xc4ac5665u,
        0xf4292244u, 0x432aff97u, 0xab9423a7u, 0xfc93a039u, 0x655b59c3u, 0x8f0ccc92u, 0xffeff47du, 0x85845dd1u,
        0x6fa87e4fu, 0xfe2ce6e0u, 0xa3014314u, 0x4e0811a1u, 0xf7537e82u, 0xbd3af235u, 0x2ad7d2bbu, 0xeb86d391u};
    // left-rotation amounts, four per round
    constexpr int S[16] = {7, 12, 17, 22, 5, 9, 14, 20, 4, 11, 16, 23, 6, 10, 15, 21};
    const uint32_t m[16] = {x, y, 0x80u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 64u, 0u};
    const uint32_t a0 = 0x67452301u, b0 = 0xefcdab89u, c0 = 0x98badcfeu, d0 = 0x10325476u;
    uint32_t a = a0, b = b0, c = c0, d = d0;
#pragma unroll
    for (int i = 0; i < 64; ++i) {
        uint32_t f;
        int g;
        if (i < 16) {
            f = (b & c) | (~b & d);
            g = i;
        } else if (i < 32) {
            f = (d & b) | (~d & c);
            g = (5 * i + 1) & 15;
        } else if (i < 48) {
            f = b ^ c ^ d;
            g = (3 * i + 5) & 15;
        } else {
            f = c ^ (b | ~d);
            g = (7 * i) & 15;
        }
        const uint32_t rot = rotl32(a + f + K[i] + m[g], S[4 * (i >> 4) + (i & 3)]);
        a = d;
        d = c;
        c = b;
        b = b + rot;
    }
    return make_uint4(a + a0, b + b0, c + c0, d + d0);
}

// Round-to-nearest-even of v to a multiple of 2^s (s >= 1), as an integer.
ORF_HD uint64_t round_bits(uint64_t v, int s)
{
    const uint64_t rem = v & ((1ull << s) - 1ull), half = 1ull << (s - 1);
    v >>= s;
    if (rem > half || (rem == half && (v & 1ull))) ++v;
    return v << s;
}

// (float)((double)l / 2^63) for l < 2^63, in integer arithmetic: l rounded to
// double's 53 significant bits, then to float's 24, each to nearest even.
ORF_HD float unit63(uint64_t l)
{
    if (l == 0ull) return 0.0f;
    int p = 63 - __builtin_clzll(l);
    if (p > 52) l = round_bits(l, p - 52);
    p = 63 - __builtin_clzll(l);
    if (p > 23) l = round_bits(l, p - 23);
    p = 63 - __builtin_clzll(l);
    const uint32_t mant = p >= 23 ? (uint32_t)(l >> (p - 23)) : (uint32_t)(l << (23 - p));
    const uint32_t e = (uint32_t)(p - 63 + 127);
    return orf_u2f((e << 23) | (mant & 0x007FFFFFu));
}

// rnd_accurate (ORF/random.h:134-142): uniform in [0, 1].
ORF_HD float rnd_accurate(Seed64& s)
{
    const uint4 h = md5_8(s.x, s.y);
    s.x = h.x;
    s.y = h.y;
    return unit63((((uint64_t)h.y << 32) | (uint64_t)h.x) & 0x7FFFFFFFFFFFFFFFull);
}

}  // namespace orf
