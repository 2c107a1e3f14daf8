// orf_dipole.h -- diffusion BSSRDFs (device; also host-callable).
//
// dipole_bssrdf: the standard dipole of Jensen et al. 2001 as
// ORF/dipoles/standard_dipole.h:8-29 writes it (real source at 3D, virtual source
// at 3D + 4AD, per RGB channel).  dirpole_bssrdf: the directional dipole of
// Frisvad et al. 2014, ORF/dipoles/directional_dipole.h:8-72.  Both return the
// diffusion term without the global coefficient (applied by the caller).
#pragma once

#include "orf_fresnel.h"
#include "orf_scene.h"

namespace orf {

ORF_HD f3 sqrt3(f3 v) { return mk3(sqrtf(v.x), sqrtf(v.y), sqrtf(v.z)); }
ORF_HD f3 exp3(f3 v) { return mk3(orf_expf(v.x), orf_expf(v.y), orf_expf(v.z)); }
ORF_HD f3 div3(f3 a, f3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }

// RoughTranslucentMaterial::computeCoefficients (ORF/RoughTranslucentMaterial.cpp:366-395)
// with an outside ior of 1, from the scaled coefficients (absorption and
// scattering already multiplied by the material's scale).
ORF_HD ScatteringGPU sss_coefficients(f3 absorption, f3 scattering, f3 meancosine, float ior, int dipole_model)
{
    ScatteringGPU p;
    p.dipole_model = dipole_model;
    p.absorption = absorption;
    p.scattering = scattering;
    p.meancosine = meancosine;
    p.relative_ior = ior;
    const f3 ior_rgb = mk3(ior);
    const f3 reducedScattering = p.scattering * (1.0f - p.meancosine);
    p.relative_ior /= (1.0f + 1.0f + 1.0f) / 3.0f;
    p.extinction = p.scattering + p.absorption;
    p.reducedExtinction = reducedScattering + p.absorption;
    p.deltaEddExtinction = p.scattering * (1.0f - p.meancosine * p.meancosine) + p.absorption;
    p.D = div3(mk3(1.0f), 3.0f * p.reducedExtinction);
    p.transport = sqrt3(div3(p.absorption, p.D));
    auto c_phi = [](float n) { return 0.25f * (1.0f - two_C1(n)); };
    auto c_e = [](float n) { return 0.5f * (1.0f - three_C2(n)); };
    p.C_phi = mk3(c_phi(ior_rgb.x), c_phi(ior_rgb.y), c_phi(ior_rgb.z));
    p.C_phi_inv = mk3(c_phi(1.0f / ior_rgb.x), c_phi(1.0f / ior_rgb.y), c_phi(1.0f / ior_rgb.z));
    p.C_E = mk3(c_e(ior_rgb.x), c_e(ior_rgb.y), c_e(ior_rgb.z));
    p.albedo = div3(p.scattering, p.extinction);
    p.reducedAlbedo = div3(reducedScattering, p.reducedExtinction);
    p.de = div3(2.131f * p.D, sqrt3(p.reducedAlbedo));
    p.A = div3(1.0f - p.C_E, 2.0f * p.C_phi);
    p.three_D = 3.0f * p.D;
    p.two_A_de = 2.0f * p.A * p.de;
    p.global_coeff = div3(mk3(1.0f), 4.0f * p.C_phi_inv) * 1.0f / (4.0f * kPi * kPi);
    p.one_over_three_ext = div3(mk3(1.0f), 3.0f * p.extinction);
    p.mean_transport = (p.transport.x + p.transport.y + p.transport.z) / 3.0f;
    return p;
}

// ORF/dipoles/standard_dipole.h:8-29.
ORF_HD f3 dipole_bssrdf(float dist, const ScatteringGPU& p)
{
    f3 real_source = p.three_D * p.three_D;
    f3 extrapolation = 4.0f * p.A * p.D;
    f3 virtual_source = extrapolation * extrapolation;
    f3 corrected_mean_free = p.three_D + extrapolation;
    f3 r_sqr = mk3(dist * dist);
    f3 d_r_sqr = r_sqr + real_source;
    f3 d_r = sqrt3(d_r_sqr);
    f3 d_v_sqr = r_sqr + virtual_source;
    f3 d_v = sqrt3(d_v_sqr);
    f3 tr_r = p.transport * d_r;
    f3 S_r = div3(p.three_D * (1.0f + tr_r), d_r_sqr * d_r);
    S_r = S_r * exp3(-tr_r);
    f3 tr_v = p.transport * d_v;
    f3 S_v = div3(corrected_mean_free * (1.0f + tr_v), d_v_sqr * d_v);
    S_v = S_v * exp3(-tr_v);
    return S_r + S_v;
}

// S_infinite / S_infinite_vec (ORF/dipoles/directional_dipole.h:8-34): the
// scalar-cosine and per-channel-cosine forms are one function here.
ORF_HD f3 S_infinite(f3 r_sqr, f3 x_dot_w12, float no_dot_w12, f3 x_dot_no, const ScatteringGPU& p)
{
    f3 r = sqrt3(r_sqr);
    f3 r_tr = p.transport * r;
    f3 r_tr_p1 = r_tr + 1.0f;
    f3 T = exp3(-r_tr);
    f3 coeff = div3(T, r * r_sqr);
    f3 first = p.C_phi * (r_sqr * p.reducedExtinction + r_tr_p1 * x_dot_w12) * 3.0f;
    f3 second = p.C_E * (p.three_D * r_tr_p1 * no_dot_w12 -
                         (r_tr_p1 + div3(p.three_D * (3.0f * r_tr_p1 + r_tr * r_tr), r_sqr) * x_dot_w12) * x_dot_no);
    return coeff * (first - second);
}

// dirpole_bssrdf (ORF/dipoles/directional_dipole.h:37-72).
ORF_HD f3 dirpole_bssrdf(f3 xi, f3 ni, f3 w12, f3 xo, f3 no, const ScatteringGPU& p)
{
    f3 x = xo - xi;
    float r_sqr = dot(x, x);
    f3 dr_sqr = mk3(r_sqr);
    float dot_x_w12 = dot(x, w12);
    f3 cos_beta = -sqrt3(div3(mk3(r_sqr - dot_x_w12 * dot_x_w12), r_sqr + p.de * p.de));
    float mu0 = -dot(no, w12);
    float edge = mu0 > 0.0f ? 1.0f : 0.0f;
    f3 D_prime = p.D * mu0 * edge + p.one_over_three_ext * (1.0f - edge);
    dr_sqr = dr_sqr + D_prime * (D_prime - 2.0f * p.de * cos_beta * edge);
    f3 t = normalize(cross(ni, x));
    f3 nistar = r_sqr < 1.0e-12f ? ni : cross(normalize(x), t);
    f3 wv = w12 - (2.0f * dot(w12, nistar)) * nistar;
    f3 xr = x - p.two_A_de.x * nistar, xg = x - p.two_A_de.y * nistar, xb = x - p.two_A_de.z * nistar;
    f3 dv_sqr = mk3(dot(xr, xr), dot(xg, xg), dot(xb, xb));
    f3 x_dot_wv = mk3(dot(xr, wv), dot(xg, wv), dot(xb, wv));
    f3 x_dot_no = mk3(dot(xr, no), dot(xg, no), dot(xb, no));
    f3 Sr = S_infinite(dr_sqr, mk3(dot_x_w12), -mu0, mk3(dot(x, no)), p);
    f3 Sv = S_infinite(dv_sqr, x_dot_wv, dot(no, wv), x_dot_no, p);
    f3 Sd = Sr - Sv;
    return mk3(fmaxf(Sd.x, 0.0f), fmaxf(Sd.y, 0.0f), fmaxf(Sd.z, 0.0f));
}

}  // namespace orf
