// orf_fresnel.h -- dielectric and conductor Fresnel reflectance in FP32.
//
// Dielectric: unpolarised reflectance as ORF/Fresnel.h:14-40.
// Conductor: the reference (ORF/Fresnel.h:78-91 on ORF/MyComplex.h) forms the
// complex cos(theta_t) = sqrt(1 - eta^2 sin^2) through polar form in double and
// returns 0.5 (|r_s|^2 + |r_p|^2).  Here the same quantities are formed in FP32
// cartesian arithmetic: eta and eta^2 come precomputed per channel, the square
// root uses the closed form sqrt(z) = (sqrt((|z|+x)/2), y / (2 sqrt((|z|+x)/2))),
// and |r|^2 is taken as |num|^2 / |den|^2 without forming the complex quotient.
// The reference's half-phase square root returns the principal root except when
// Re z < 0 and Im z < 0, where it returns its negative; that case is kept.
#pragma once

#include "orf_vec.h"

namespace orf {

struct cf { float re, im; };

// Per-channel conductor constants: eta = n1/ior and eta^2.
struct ConductorEta {
    cf eta[3];
    cf eta2[3];
    int gray;  // all three channels equal: evaluate once
    cf ieta[3];   // 1 / eta (= ior): a microsurface walk event met from below
    cf ieta2[3];
};

ORF_HD cf csqrt_ref(float zr, float zi)
{
    float m = sqrtf(zr * zr + zi * zi);
    cf r;
    if (zr >= 0.0f) {
        float t = sqrtf(0.5f * (m + zr));
        r.re = t;
        r.im = t > 0.0f ? zi / (2.0f * t) : 0.0f;
    } else {
        // principal root is (|zi| / 2t, sign(zi) t); the reference's branch
        // negates it for zi < 0, leaving (-|zi| / 2t, t) for both signs
        float t = sqrtf(0.5f * (m - zr));
        r.im = t;
        r.re = zi / (2.0f * t);
    }
    return r;
}

ORF_HD float fresnel_conductor_channel(float cos_i, float sin_i, cf eta, cf eta2)
{
    float s2 = sin_i * sin_i;
    cf ct = csqrt_ref(1.0f - eta2.re * s2, -(eta2.im * s2));
    float ecr = eta.re * cos_i, eci = eta.im * cos_i;
    float nr = ecr - ct.re, ni = eci - ct.im;
    float dr = ecr + ct.re, di = eci + ct.im;
    float r_s = (nr * nr + ni * ni) / (dr * dr + di * di);
    float er = eta.re * ct.re - eta.im * ct.im;
    float ei = eta.re * ct.im + eta.im * ct.re;
    float pn = (cos_i - er) * (cos_i - er) + ei * ei;
    float pd = (cos_i + er) * (cos_i + er) + ei * ei;
    float r_p = pn / pd;
    return 0.5f * (r_s + r_p);
}

ORF_HD f3 fresnel_conductor(f3 w_i, f3 w_m, const ConductorEta& e)
{
    float c = dot(w_i, w_m);
    float s = sqrtf(1.0f - c * c);
    float r = fresnel_conductor_channel(c, s, e.eta[0], e.eta2[0]);
    if (e.gray) return mk3(r);
    return mk3(r, fresnel_conductor_channel(c, s, e.eta[1], e.eta2[1]),
               fresnel_conductor_channel(c, s, e.eta[2], e.eta2[2]));
}

// fresnel_conductor with 1 / eta (the walk's events met from below).
ORF_HD f3 fresnel_conductor_inv(f3 w_i, f3 w_m, const ConductorEta& e)
{
    float c = dot(w_i, w_m);
    float s = sqrtf(1.0f - c * c);
    float r = fresnel_conductor_channel(c, s, e.ieta[0], e.ieta2[0]);
    if (e.gray) return mk3(r);
    return mk3(r, fresnel_conductor_channel(c, s, e.ieta[1], e.ieta2[1]),
               fresnel_conductor_channel(c, s, e.ieta[2], e.ieta2[2]));
}

// Unpolarised dielectric reflectance for relative IOR ior1/ior2.
ORF_HD float fresnel_dielectric(float cos_theta, float eta)
{
    float st2 = eta * eta * (1.0f - cos_theta * cos_theta);
    if (st2 >= 1.0f) return 1.0f;
    float ct = sqrtf(1.0f - st2);
    float rs = (eta * cos_theta - ct) / (eta * cos_theta + ct);
    float rp = (cos_theta - eta * ct) / (cos_theta + eta * ct);
    return (rs * rs + rp * rp) * 0.5f;
}

// two_C1(n) (ORF/Fresnel.h:93-101): twice the first Fresnel moment, the
// published polynomial fits of d'Eon & Irving 2011 (one per side of n = 1).
ORF_HD float two_C1(float n)
{
    if (n >= 1.0f) return -9.23372f + n * (22.2272f + n * (-20.9292f + n * (10.2291f + n * (-2.54396f + n * 0.254913f))));
    return 0.919317f + n * (-3.4793f + n * (6.75335f + n * (-7.80989f + n * (4.98554f - n * 1.36881f))));
}

// three_C2(n) (ORF/Fresnel.h:103-114): three times the second Fresnel moment,
// the same authors' fits.
ORF_HD float three_C2(float n)
{
    float r;
    if (n >= 1.0f) {
        r = -1641.1f + n * (1213.67f + n * (-568.556f + n * (164.798f + n * (-27.0181f + n * 1.91826f))));
        r += (((135.926f / n) - 656.175f) / n + 1376.53f) / n;
    } else {
        r = 0.828421f + n * (-2.62051f + n * (3.36231f + n * (-1.95284f + n * (0.236494f + n * 0.145787f))));
    }
    return r;
}

}  // namespace orf
