// orf_runtime.cpp -- library-wide C-ABI pieces: errors, version, workload generators.
#include <cmath>
#include <cstdint>
#include <string>

#include "orf_common.h"
#include "orf_render.h"
#include "orf_rng.h"

namespace orf {

// eta = 1 / ior as the reference forms it (operator/(float, MyComplex),
// ORF/MyComplex.h:155-160: conj(ior) / |ior|^2 with |ior| rounded to float).
void conductor_eta_host(const double ior_re[3], const double ior_im[3], ConductorEta& e)
{
    for (int c = 0; c < 3; ++c) {
        float r = (float)ior_re[c], i = (float)ior_im[c];
        float ab = sqrtf(r * r + i * i);
        double den = (double)powf(ab, 2.0f);
        double er = ior_re[c] * (1.0 / den);
        double ei = -ior_im[c] * (1.0 / den);
        e.eta[c] = cf{(float)er, (float)ei};
        e.eta2[c] = cf{(float)(er * er - ei * ei), (float)(2.0 * er * ei)};
        e.ieta[c] = cf{(float)ior_re[c], (float)ior_im[c]};
        e.ieta2[c] = cf{(float)(ior_re[c] * ior_re[c] - ior_im[c] * ior_im[c]), (float)(2.0 * ior_re[c] * ior_im[c])};
    }
    e.gray = (ior_re[0] == ior_re[1] && ior_re[1] == ior_re[2] && ior_im[0] == ior_im[1] &&
              ior_im[1] == ior_im[2]);
}

// A dielectric interface's real relative index n12 as a conductor's eta
// (CU/rough_diffuse_shader.cu:132-133, :254-255 pass MyComplex{n1_over_n2, 0});
// 1 / eta formed as operator/(float, MyComplex) forms it.
void dielectric_eta_host(float n12, ConductorEta& e)
{
    const float ab = sqrtf(n12 * n12);
    const double ie = (double)n12 * (1.0 / (double)powf(ab, 2.0f));
    for (int c = 0; c < 3; ++c) {
        e.eta[c] = cf{n12, 0.0f};
        e.eta2[c] = cf{(float)((double)n12 * (double)n12), 0.0f};
        e.ieta[c] = cf{(float)ie, 0.0f};
        e.ieta2[c] = cf{(float)(ie * ie), 0.0f};
    }
    e.gray = 1;
}

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

// PCG32 (O'Neill 2014, pcg-c-basic): multiplier 6364136223846793005, XSH-RR output.
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    Pcg32(uint64_t initstate, uint64_t initseq)
    {
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    uint32_t next()
    {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
};

}  // namespace orf

extern "C" const char* orf_last_error(void) { return orf::g_last_error.c_str(); }

extern "C" const char* orf_version(void) { return "orf_rt 0.1 gfx950"; }

extern "C" orf_status orf_gen_bsdf_queries(int64_t n, uint64_t seed, uint64_t seq, float* wo_x,
                                           float* wo_y, float* wo_z, float* alpha, float* u1,
                                           float* u2)
{
    ORF_API_BEGIN
    ORF_REQUIRE(n >= 0, "orf_gen_bsdf_queries: n < 0");
    ORF_REQUIRE(n == 0 || (wo_x && wo_y && wo_z && alpha && u1 && u2),
                "orf_gen_bsdf_queries: null array");
    orf::Pcg32 rng(seed, seq);
    const float two_pi = 6.28318530717958647692f;
    for (int64_t i = 0; i < n; ++i) {
        float u[5];
        for (float& v : u) v = (float)(rng.next() >> 8) * (1.0f / 16777216.0f);
        float c = u[0];
        float s = std::sqrt(std::fmax(0.0f, 1.0f - c * c));
        float phi = two_pi * u[1];
        wo_x[i] = s * std::cos(phi);
        wo_y[i] = s * std::sin(phi);
        wo_z[i] = c;
        alpha[i] = 0.05f + 0.95f * u[2];
        u1[i] = u[3];
        u2[i] = u[4];
    }
    ORF_API_END
}

extern "C" orf_status orf_rng_accurate_stream(uint32_t x, uint32_t y, int n, uint32_t* states, float* values)
{
    ORF_API_BEGIN
    ORF_REQUIRE(n >= 0 && (n == 0 || (states && values)), "orf_rng_accurate_stream: bad arguments");
    orf::Seed64 s{x, y};
    for (int i = 0; i < n; ++i) {
        values[i] = orf::rnd_accurate(s);
        states[2 * i] = s.x;
        states[2 * i + 1] = s.y;
    }
    ORF_API_END
}

extern "C" orf_status orf_rng_unit63(const uint64_t* l, int n, float* out)
{
    ORF_API_BEGIN
    ORF_REQUIRE(n >= 0 && (n == 0 || (l && out)), "orf_rng_unit63: bad arguments");
    for (int i = 0; i < n; ++i) {
        ORF_REQUIRE(l[i] < (1ull << 63), "orf_rng_unit63: value of 64 bits");
        out[i] = orf::unit63(l[i]);
    }
    ORF_API_END
}
