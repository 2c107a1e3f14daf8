// orf_sss_grad.h -- derivatives of the diffusion term with respect to the
// translucent material's absorption and scattering coefficients (device; also
// host-callable).  No counterpart in the reference.
//
// Forward mode: a value with two tangents, d/d sigma_a and d/d sigma_s of one
// colour channel (channels are independent: every coefficient of channel c
// depends on that channel's sigma_a, sigma_s only).  The coefficient formulas
// are RoughTranslucentMaterial::computeCoefficients'
// (ORF/RoughTranslucentMaterial.cpp:366-395), the dipoles ORF/dipoles/
// standard_dipole.h:8-29 and directional_dipole.h:8-72, all restated per
// channel on this type; C_phi, C_E, A and the global coefficient depend on the
// index of refraction only.
#pragma once

#include "orf_dipole.h"

namespace orf {

struct Dual2 {
    float v, a, s;  // value, d/d sigma_a, d/d sigma_s
};

ORF_HD Dual2 dc(float v) { return Dual2{v, 0.0f, 0.0f}; }
ORF_HD Dual2 operator+(Dual2 x, Dual2 y) { return Dual2{x.v + y.v, x.a + y.a, x.s + y.s}; }
ORF_HD Dual2 operator-(Dual2 x, Dual2 y) { return Dual2{x.v - y.v, x.a - y.a, x.s - y.s}; }
ORF_HD Dual2 operator-(Dual2 x) { return Dual2{-x.v, -x.a, -x.s}; }
ORF_HD Dual2 operator*(Dual2 x, Dual2 y) { return Dual2{x.v * y.v, x.a * y.v + x.v * y.a, x.s * y.v + x.v * y.s}; }
ORF_HD Dual2 operator*(Dual2 x, float k) { return Dual2{x.v * k, x.a * k, x.s * k}; }
ORF_HD Dual2 operator*(float k, Dual2 x) { return x * k; }
ORF_HD Dual2 operator+(Dual2 x, float k) { return Dual2{x.v + k, x.a, x.s}; }
ORF_HD Dual2 operator/(Dual2 x, Dual2 y)
{
    const float q = x.v / y.v;
    return Dual2{q, (x.a - q * y.a) / y.v, (x.s - q * y.s) / y.v};
}
ORF_HD Dual2 operator/(float k, Dual2 y) { return dc(k) / y; }
ORF_HD Dual2 dsqrt(Dual2 x)
{
    const float r = sqrtf(x.v);
    const float h = r > 0.0f ? 0.5f / r : 0.0f;
    return Dual2{r, x.a * h, x.s * h};
}
ORF_HD Dual2 dexp(Dual2 x)
{
    const float e = orf_expf(x.v);
    return Dual2{e, x.a * e, x.s * e};
}

// One channel's derived coefficients as functions of its (scaled) sigma_a, sigma_s.
struct SssCoefD {
    Dual2 reducedExtinction, D, transport, de, three_D, two_A_de, one_over_three_ext;
    float C_phi, C_E, A;
};

ORF_HD SssCoefD sss_coef_channel(float sigma_a, float sigma_s, float g, float ior)
{
    SssCoefD p;
    const Dual2 sa{sigma_a, 1.0f, 0.0f}, ss{sigma_s, 0.0f, 1.0f};
    const Dual2 red = ss * (1.0f - g);
    const Dual2 ext = ss + sa;
    p.reducedExtinction = red + sa;
    p.D = 1.0f / (3.0f * p.reducedExtinction);
    p.transport = dsqrt(sa / p.D);
    p.C_phi = 0.25f * (1.0f - two_C1(ior));
    p.C_E = 0.5f * (1.0f - three_C2(ior));
    p.A = (1.0f - p.C_E) / (2.0f * p.C_phi);
    const Dual2 reducedAlbedo = red / p.reducedExtinction;
    p.de = (2.131f * p.D) / dsqrt(reducedAlbedo);
    p.three_D = 3.0f * p.D;
    p.two_A_de = (2.0f * p.A) * p.de;
    p.one_over_three_ext = 1.0f / (3.0f * ext);
    return p;
}

// The standard dipole of one channel (ORF/dipoles/standard_dipole.h:8-29).
ORF_HD Dual2 dipole_channel(float dist, const SssCoefD& p)
{
    const Dual2 real_source = p.three_D * p.three_D;
    const Dual2 extrapolation = (4.0f * p.A) * p.D;
    const Dual2 virtual_source = extrapolation * extrapolation;
    const Dual2 corrected_mean_free = p.three_D + extrapolation;
    const float r_sqr = dist * dist;
    const Dual2 d_r_sqr = real_source + r_sqr, d_v_sqr = virtual_source + r_sqr;
    const Dual2 d_r = dsqrt(d_r_sqr), d_v = dsqrt(d_v_sqr);
    const Dual2 tr_r = p.transport * d_r, tr_v = p.transport * d_v;
    const Dual2 S_r = p.three_D * (tr_r + 1.0f) / (d_r_sqr * d_r) * dexp(-tr_r);
    const Dual2 S_v = corrected_mean_free * (tr_v + 1.0f) / (d_v_sqr * d_v) * dexp(-tr_v);
    return S_r + S_v;
}

// S_infinite of one channel (ORF/dipoles/directional_dipole.h:8-20).
ORF_HD Dual2 S_infinite_channel(Dual2 r_sqr, Dual2 x_dot_w12, float no_dot_w12, Dual2 x_dot_no, const SssCoefD& p,
                                Dual2 reducedExtinction)
{
    const Dual2 r = dsqrt(r_sqr);
    const Dual2 r_tr = p.transport * r;
    const Dual2 r_tr_p1 = r_tr + 1.0f;
    const Dual2 coeff = dexp(-r_tr) / (r * r_sqr);
    const Dual2 first = (r_sqr * reducedExtinction + r_tr_p1 * x_dot_w12) * (p.C_phi * 3.0f);
    const Dual2 second =
        p.C_E * (p.three_D * r_tr_p1 * no_dot_w12 - (r_tr_p1 + p.three_D * (3.0f * r_tr_p1 + r_tr * r_tr) / r_sqr * x_dot_w12) * x_dot_no);
    return coeff * (first - second);
}

// The directional dipole of one channel (ORF/dipoles/directional_dipole.h:37-72),
// clamped at zero as the reference clamps it.
ORF_HD Dual2 dirpole_channel(f3 xi, f3 ni, f3 w12, f3 xo, f3 no, const SssCoefD& p)
{
    const f3 x = xo - xi;
    const float r_sqr = dot(x, x);
    const float dot_x_w12 = dot(x, w12);
    const Dual2 cos_beta = -dsqrt(dc(r_sqr - dot_x_w12 * dot_x_w12) / (p.de * p.de + r_sqr));
    const float mu0 = -dot(no, w12);
    const float edge = mu0 > 0.0f ? 1.0f : 0.0f;
    const Dual2 D_prime = p.D * (mu0 * edge) + p.one_over_three_ext * (1.0f - edge);
    const Dual2 dr_sqr = D_prime * (D_prime - (2.0f * p.de) * cos_beta * edge) + r_sqr;
    const f3 t = normalize(cross(ni, x));
    const f3 nistar = r_sqr < 1.0e-12f ? ni : cross(normalize(x), t);
    const f3 wv = w12 - (2.0f * dot(w12, nistar)) * nistar;
    // x - two_A_de nistar, with two_A_de a dual: the squared length and the two cosines
    const float xn = dot(x, nistar), nn = dot(nistar, nistar);
    const Dual2 k = p.two_A_de;
    const Dual2 dv_sqr = (k * k) * nn - (2.0f * xn) * k + r_sqr;
    const Dual2 x_dot_wv = dc(dot(x, wv)) - k * dot(nistar, wv);
    const Dual2 x_dot_no = dc(dot(x, no)) - k * dot(nistar, no);
    const Dual2 Sr = S_infinite_channel(dr_sqr, dc(dot_x_w12), -mu0, dc(dot(x, no)), p, p.reducedExtinction);
    const Dual2 Sv = S_infinite_channel(dv_sqr, x_dot_wv, dot(no, wv), x_dot_no, p, p.reducedExtinction);
    const Dual2 Sd = Sr - Sv;
    return Sd.v > 0.0f ? Sd : dc(0.0f);
}

// What a translucent vertex of the differentiated material contributes to the
// coefficient gradients: d log(extinction weight) (inside vertices) and the
// derivative of the radiance it adds itself (the diffusion gather), per channel,
// with respect to the scaled sigma_a and sigma_s.
struct SssGrad {
    bool beam;          // an extinction weight exp(-t sigma_t) / p was applied
    f3 dlog_a, dlog_s;  // its d log / d sigma_a, d sigma_s
    f3 dL_a, dL_s;      // d(radiance added here) / d sigma_a, d sigma_s
};

ORF_HD void sss_grad_clear(SssGrad& g)
{
    g.beam = false;
    g.dlog_a = g.dlog_s = g.dL_a = g.dL_s = mk3(0.0f);
}

// The extinction weight exp(-t sigma_t) / p inside the medium, its probability
// p held fixed: d log / d sigma = -t d sigma_t / d sigma, where sigma_t is the
// extinction sigma_a + sigma_s (standard dipole) or the delta-Eddington
// extinction sigma_a + sigma_s (1 - g^2) (directional), as the programs pick it.
ORF_HD void sss_beam_grad(SssGrad& g, const ScatteringGPU& P, float t)
{
    g.beam = true;
    g.dlog_a = mk3(-t);
    g.dlog_s = P.dipole_model == 1 ? (-t) * (1.0f - P.meancosine * P.meancosine) : mk3(-t);
}

}  // namespace orf
