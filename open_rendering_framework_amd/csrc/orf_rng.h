// orf_rng.h -- per-path random numbers (device and host).
//
// The render path keeps one 32-bit state per path and advances it with 16
// TEA rounds per draw, exactly the stream of ORF/random.h:25-40 (tea<N>) and
// :117-126 (tea_hash / rnd_tea); the pixel seed is tea<16>(W*y + x, frame)
// (CU/path_tracer.cu:48).  Bit-exactness of this stream is what lets the
// oracle and the kernels follow the same paths.
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

}  // namespace orf
