/*
 * oracle_math.c -- CPU restatement of the reference's device math (TEST INFRASTRUCTURE).
 *
 * Follows ORF/random.h, ORF/helpers.h, ORF/sampler.h, ORF/MicrofacetGGX.h,
 * ORF/Microfacet.h, ORF/Fresnel.h, ORF/MyComplex.h and CU/metallic_shader.cu.
 * Only tests, the smoke check and the CPU-baseline leg of the benchmark use it.
 *
 * Evaluation order of every expression follows the reference so that FP32
 * rounding matches the reference's host-compiled arithmetic; the product's
 * kernels are compiled without FMA contraction for the same reason.
 */
#include <tgmath.h>
#include <math.h>
#include <string.h>
#include "orf_oracle.h"
#include "oracle_fmath.h"
#include "../open_rendering_framework_amd/csrc/orf_math.h"

#define PI_R ((R)3.14159265358979323846)

/* ------------------------------------------------------------------ RNG */

/* tea<N> (ORF/random.h:25-40): TEA encryption rounds with the reference's key. */
uint32_t orc_tea_key(uint32_t v0, uint32_t v1, uint32_t rounds, const uint32_t key[4])
{
    uint32_t s0 = 0;
    for (uint32_t n = 0; n < rounds; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + key[0]) ^ (v1 + s0) ^ ((v1 >> 5) + key[1]);
        v1 += ((v0 << 4) + key[2]) ^ (v0 + s0) ^ ((v0 >> 5) + key[3]);
    }
    return v0;
}

static const uint32_t REF_TEA_KEY[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};

uint32_t orc_tea16(uint32_t v0, uint32_t v1) { return orc_tea_key(v0, v1, 16, REF_TEA_KEY); }

/* tea_hash / rnd_tea (ORF/random.h:117-126).  The value is formed in float as
 * in the reference (it can round up to exactly 1.0f); the FP64 oracle promotes
 * that same float so both oracles consume identical random numbers. */
R orc_rnd_tea(uint32_t *prev)
{
    *prev = orc_tea16(*prev, 100u);
    uint32_t h = *prev & 0x7FFFFFFFu;
    float u = (float)h / (float)0x80000000u;
    return (R)u;
}

/* lcg / rnd (ORF/random.h:43-61): Numerical Recipes LCG, 24-bit output. */
uint32_t orc_lcg(uint32_t *prev)
{
    *prev = 1664525u * *prev + 1013904223u;
    return *prev & 0x00FFFFFFu;
}

R orc_rnd(uint32_t *prev) { return (R)((float)orc_lcg(prev) / (float)0x01000000); }

/* __md5Hash (ORF/md5.h:58-225): MD5 (RFC 1321) of a message of at most 55
 * bytes, so one padded 512-bit block: the message bytes packed little-endian
 * into words, 0x80 after them, the bit length in word 14.  The additive
 * constants are formed from their definition, floor(|sin(i + 1)| 2^32), and the
 * four rounds use F, G, H, I with the RFC's shifts and message-word orders. */
void orc_md5(const uint8_t* msg, uint32_t len, uint32_t out[4])
{
    static uint32_t K[64];
    static int have_k = 0;
    if (!have_k) {
        for (int i = 0; i < 64; i++) K[i] = (uint32_t)floor(fabs(sin((double)(i + 1))) * 4294967296.0);
        have_k = 1;
    }
    static const int SH[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
    uint32_t w[16];
    memset(w, 0, sizeof(w));
    uint32_t i;
    for (i = 0; i < len && i < 55; i++) w[i / 4] |= (uint32_t)msg[i] << ((i % 4) * 8);
    w[i / 4] |= 0x80u << ((i % 4) * 8);
    w[14] = len * 8u;
    const uint32_t init[4] = {0x67452301u, 0xEFCDAB89u, 0x98BADCFEu, 0x10325476u};
    uint32_t a = init[0], b = init[1], c = init[2], d = init[3];
    for (int r = 0; r < 4; r++) {
        for (int k = 0; k < 16; k++) {
            int step = 16 * r + k, idx;
            uint32_t f;
            switch (r) {
            case 0: f = (b & c) | (~b & d); idx = k; break;
            case 1: f = (b & d) | (c & ~d); idx = (1 + 5 * k) % 16; break;
            case 2: f = b ^ c ^ d; idx = (5 + 3 * k) % 16; break;
            default: f = c ^ (b | ~d); idx = (7 * k) % 16; break;
            }
            uint32_t v = a + f + w[idx] + K[step];
            int sh = SH[r][k % 4];
            v = (v << sh) | (v >> (32 - sh));
            a = d;
            d = c;
            c = b;
            b = b + v;
        }
    }
    out[0] = a + init[0];
    out[1] = b + init[1];
    out[2] = c + init[2];
    out[3] = d + init[3];
}

/* rnd_accurate (ORF/random.h:128-142): the state (two words, Seed64's uint2)
 * hashed as 8 bytes (rand_md5 with scrambler 0, ORF/md5.h:242-249), the state
 * replaced by the first two digest words, the value (low 63 bits of the 64-bit
 * state) / 2^63 in double, then rounded to float. */
R orc_rnd_accurate(uint32_t s[2])
{
    uint8_t m[8];
    for (int k = 0; k < 4; k++) {
        m[k] = (uint8_t)(s[0] >> (8 * k));
        m[4 + k] = (uint8_t)(s[1] >> (8 * k));
    }
    uint32_t h[4];
    orc_md5(m, 8, h);
    s[0] = h[0];
    s[1] = h[1];
    uint64_t l = ((uint64_t)h[1] << 32 | (uint64_t)h[0]) & 0x7FFFFFFFFFFFFFFFull;
    double val = (double)l / 9223372036854775808.0;
    return (R)(float)val;
}

/* PCG32, pcg-c-basic (O'Neill 2014): published algorithm, used only to
 * generate the synthetic operator queries. */
uint32_t orc_pcg32_next(orc_pcg32 *r)
{
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}

void orc_pcg32_seed(orc_pcg32 *r, uint64_t initstate, uint64_t initseq)
{
    r->state = 0u;
    r->inc = (initseq << 1u) | 1u;
    orc_pcg32_next(r);
    r->state += initstate;
    orc_pcg32_next(r);
}

/* ------------------------------------------------------------ vectors */

/* optixu normalize: multiply by the reciprocal length. */
v3 orc_normalize(v3 v)
{
    R inv = (R)1 / sqrt(vdot(v, v));
    return vscale(v, inv);
}

/* faceforward(n, i, nref) = n * sign(dot(i, nref)) (OptiX convention). */
v3 orc_faceforward(v3 n, v3 i, v3 nref)
{
    return vscale(n, copysign((R)1, vdot(i, nref)));
}

/* create_onb, two-output form (ORF/helpers.h:107-119). */
void orc_create_onb(v3 n, v3 *U, v3 *V)
{
    v3 u = vcross(n, V3(0, 1, 0));
    if (vdot(u, u) < (R)1e-3f)
        u = vcross(n, V3(1, 0, 0));
    u = orc_normalize(u);
    *U = u;
    *V = vcross(n, u);
}

/* transformToLocal / transformToWorld (ORF/helpers.h:284-299). */
v3 orc_to_local(v3 w, v3 n)
{
    v3 u, v;
    orc_create_onb(n, &u, &v);
    return orc_normalize(V3(vdot(w, u), vdot(w, v), vdot(w, n)));
}

v3 orc_to_world(v3 w, v3 n)
{
    v3 u, v;
    orc_create_onb(n, &u, &v);
    return orc_normalize(vadd(vadd(vscale(u, w.x), vscale(v, w.y)), vscale(n, w.z)));
}

/* ------------------------------------------------------------ sampler.h */

/* rotate_to_normal (ORF/sampler.h:12-24, Frisvad 2012). */
void orc_rotate_to_normal(v3 n, v3 *v)
{
    if (n.z < (R)-0.999999f) {
        *v = V3(-v->y, -v->x, -v->z);
        return;
    }
    R a = (R)1 / ((R)1 + n.z);
    R b = -n.x * n.y * a;
    v3 c0 = V3((R)1 - n.x * n.x * a, b, -n.x);
    v3 c1 = V3(b, (R)1 - n.y * n.y * a, -n.y);
    *v = vadd(vadd(vscale(c0, v->x), vscale(c1, v->y)), vscale(n, v->z));
}

/* sample_cosine_weighted (ORF/sampler.h:51-64) with spherical_direction (:29-34). */
v3 orc_sample_cosine_weighted(v3 n, uint32_t *t)
{
    R cos_theta = sqrt(orc_rnd_tea(t));
    R phi = (R)2 * PI_R * orc_rnd_tea(t);
    R sin_theta = sqrt((R)1 - cos_theta * cos_theta);
    v3 v = V3(sin_theta * F_COS(phi), sin_theta * F_SIN(phi), cos_theta);
    orc_rotate_to_normal(n, &v);
    return v;
}

/* sample_barycentric (ORF/sampler.h:103-116). */
v3 orc_sample_barycentric(uint32_t *t)
{
    R sqrt_xi1 = sqrt(orc_rnd_tea(t));
    R xi2 = orc_rnd_tea(t);
    return V3((R)1 - sqrt_xi1, ((R)1 - xi2) * sqrt_xi1, xi2 * sqrt_xi1);
}

/* ------------------------------------------------------------ GGX */

/* ggx_alpha_i (ORF/MicrofacetGGX.h:9-17). */
R orc_ggx_alpha_i(v3 wi, R ax, R ay)
{
    R inv_sin2 = (R)1 / ((R)1 - wi.z * wi.z);
    if (wi.z == (R)1 || wi.z == (R)-1)
        inv_sin2 = 0;
    R cos_phi2 = wi.x * wi.x * inv_sin2;
    R sin_phi2 = wi.y * wi.y * inv_sin2;
    return sqrt(cos_phi2 * ax * ax + sin_phi2 * ay * ay);
}

/* ggx_G1(cos, alpha^2) (ORF/MicrofacetGGX.h:54-60): Smith G1 closed form. */
R orc_ggx_G1_cos(R cos_theta, R alpha_sqr)
{
    R c2 = cos_theta * cos_theta;
    R t2 = ((R)1 - c2) / c2;
    return (R)2 / ((R)1 + sqrt((R)1 + alpha_sqr * t2));
}

/* ggx_G1(v, m, n, ax, ay) (ORF/MicrofacetGGX.h:63-74). */
R orc_ggx_G1(v3 v, v3 m, v3 n, R ax, R ay)
{
    R alpha = orc_ggx_alpha_i(orc_normalize(orc_to_local(v, n)), ax, ay);
    R a2 = alpha * alpha;
    R cos_v = vdot(v, n);
    if (cos_v * vdot(v, m) <= 0)
        return 0;
    return orc_ggx_G1_cos(cos_v, a2);
}

/* ggx_projected_area (ORF/MicrofacetGGX.h:33-45): cos(theta) (1 + Lambda). */
R orc_ggx_projected_area(v3 wi, R ax, R ay)
{
    if (wi.z > (R)0.9999f)
        return 1;
    if (wi.z < (R)-0.9999f)
        return 0;
    R alpha = orc_ggx_alpha_i(wi, ax, ay);
    R theta = F_ACOS(wi.z);
    R s = F_SIN(theta);
    return (R)0.5f * (wi.z + sqrt(wi.z * wi.z + s * s * alpha * alpha));
}

/* ggx_slope_pdf + ggx_eval_NDF (ORF/MicrofacetGGX.h:48-53, 219-231). */
R orc_ggx_eval_ndf(v3 wm, R ax, R ay)
{
    if (wm.z <= 0)
        return 0;
    R sx = -wm.x / wm.z;
    R sy = -wm.y / wm.z;
    R tmp = (R)1 + sx * sx / (ax * ax) + sy * sy / (ay * ay);
    R p22 = (R)1 / (PI_R * ax * ay) / (tmp * tmp);
    return p22 / (wm.z * wm.z * wm.z * wm.z);
}

/* ggx_eval_VNDF (ORF/MicrofacetGGX.h:234-251): D_wi(m) = <wi,m> D(m) / proj_area. */
R orc_ggx_eval_vndf(v3 wi, v3 wm, R ax, R ay)
{
    if (wm.z <= 0)
        return 0;
    R area = orc_ggx_projected_area(wi, ax, ay);
    if (area == 0)
        return 0;
    R c = (R)1 / area;
    R D = orc_ggx_eval_ndf(wm, ax, ay);
    R d = vdot(wi, wm);
    return c * (d > 0 ? d : 0) * D;
}

/* ggx_sample_P22_11 (ORF/MicrofacetGGX.h:142-187): visible-slope sampling of
 * the unit-roughness distribution (Heitz & d'Eon 2014, rational fit at :184). */
v2 orc_ggx_sample_p22_11(R theta_i, R z1, R z2)
{
    v2 s;
    if (theta_i < (R)0.0001f) {
        R r = sqrt(z1 / ((R)1 - z1));
        R phi = PI_R * (R)2 * z2;
        s.x = r * F_COS(phi);
        s.y = r * F_SIN(phi);
        return s;
    }
    R sin_i = F_SIN(theta_i);
    R cos_i = F_COS(theta_i);
    if (!(cos_i >= (R)1e-7f)) cos_i = (R)1e-7f;  /* fmaxf(cos, 1e-7) */
    R tan_i = sin_i / cos_i;
    R proj_area = (R)0.5f * (cos_i + (R)1);
    if (proj_area < (R)0.0001f || proj_area != proj_area) {
        s.x = 0; s.y = 0;
        return s;
    }
    R c = (R)1 / proj_area;
    R A = (R)2 * z1 / cos_i / c - (R)1;
    R B = tan_i;
    R tmp = (R)1 / (A * A - (R)1);
    R disc = B * B * tmp * tmp - (A * A - B * B) * tmp;
    R D = sqrt(disc > 0 ? disc : 0);
    R sx1 = B * tmp - D;
    R sx2 = B * tmp + D;
    s.x = (A < 0 || sx2 > (R)1 / tan_i) ? sx1 : sx2;
    R S, zz2;
    if (z2 > (R)0.5f) {
        S = 1;
        zz2 = (R)2 * (z2 - (R)0.5f);
    } else {
        S = -1;
        zz2 = (R)2 * ((R)0.5f - z2);
    }
    R z = (zz2 * (zz2 * (zz2 * (R)0.27385f - (R)0.73369f) + (R)0.46341f)) /
          (zz2 * (zz2 * (zz2 * (R)0.093073f + (R)0.309420f) - (R)1.0f) + (R)0.597999f);
    s.y = S * z * sqrt((R)1 + s.x * s.x);
    return s;
}

/* ggx_sample_VNDF (ORF/MicrofacetGGX.h:190-207). */
v3 orc_ggx_sample_vndf(v3 wi, R ax, R ay, R z1, R z2)
{
    v3 wi11 = orc_normalize(V3(ax * wi.x, ay * wi.y, wi.z));
    v2 s11 = orc_ggx_sample_p22_11(F_ACOS(wi11.z), z1, z2);
    R phi = F_ATAN2(wi11.y, wi11.x);
    R cp = F_COS(phi), sp = F_SIN(phi);
    R sx = cp * s11.x - sp * s11.y;
    R sy = sp * s11.x + cp * s11.y;
    sx *= ax;
    sy *= ay;
    if (!isfinite(sx) || sx != sx) {
        if (wi.z > 0)
            return V3(0, 0, 1);
        return orc_normalize(V3(wi.x, wi.y, 0));
    }
    return orc_normalize(V3(-sx, -sy, (R)1));
}

/* ------------------------------------------------------------ complex Fresnel */

/* MyComplex helpers (ORF/MyComplex.h).  The reference forms |c| and arg(c)
 * in float (abs :18-23, phase :169-185); R reproduces that truncation in the
 * FP32 oracle and removes it in the FP64 oracle. */
static R c_abs(cplx c)
{
    R r = (R)c.re, i = (R)c.im;
    return (R)sqrt((double)(r * r + i * i));
}

static R c_phase(cplx c)
{
    double r = c.re, i = c.im, th;
    const double pi_ref = (double)PI_R;
    if (r == 0 && i > 0) th = pi_ref / 2;
    else if (r == 0 && i < 0) th = -pi_ref / 2;
    else if (r == 0 && i == 0) th = 0;
    else th = atan(i / r);
    if (r < 0) th += pi_ref;
    return (R)th;
}

static cplx C(double re, double im) { cplx c = {re, im}; return c; }
static cplx c_add(cplx a, cplx b) { return C(a.re + b.re, a.im + b.im); }
static cplx c_sub(cplx a, cplx b) { return C(a.re - b.re, a.im - b.im); }
static cplx c_mul(cplx a, cplx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static cplx c_scale(cplx a, double s) { return C(a.re * s, a.im * s); }

/* operator/(MyComplex, MyComplex) (ORF/MyComplex.h:148-153). */
static cplx c_div(cplx a, cplx b)
{
    cplx num = c_mul(a, C(b.re, -b.im));
    R ab = c_abs(b);
    double den = (double)(R)pow(ab, (R)2);
    return c_scale(num, 1.0 / den);
}

/* operator/(float, MyComplex) (ORF/MyComplex.h:155-160). */
cplx orc_cdiv_fc(R c1, cplx b)
{
    cplx num = C((double)c1 * b.re, (double)c1 * -b.im);
    R ab = c_abs(b);
    double den = (double)(R)pow(ab, (R)2);
    return c_scale(num, 1.0 / den);
}

/* pow(MyComplex, int) (ORF/MyComplex.h:187-193). */
static cplx c_powi(cplx c, int e)
{
    double ab = pow((double)c_abs(c), (double)e);
    double th = (double)(R)(c_phase(c) * (R)e);
    return C(ab * cos(th), ab * sin(th));
}

/* sqrt(MyComplex) (ORF/MyComplex.h:195-202): half the phase in (-pi/2, 3pi/2]. */
static cplx c_sqrt(cplx c)
{
    double ab = sqrt((double)c_abs(c));
    double th = (double)(R)(c_phase(c) / (R)2);
    return C(ab * cos(th), ab * sin(th));
}

static R fresnel_conductor_1(R cos_i, R sin_i, cplx eta)
{
    cplx sin_t = c_scale(eta, (double)sin_i);
    cplx p = c_powi(sin_t, 2);
    cplx cos_t = c_sqrt(C((double)(R)1.0f - p.re, -p.im));
    cplx e_ci = c_scale(eta, (double)cos_i);
    cplx r_ort = c_div(c_sub(e_ci, cos_t), c_add(e_ci, cos_t));
    cplx e_ct = c_mul(eta, cos_t);
    cplx r_par = c_div(C((double)cos_i - e_ct.re, -e_ct.im), C((double)cos_i + e_ct.re, e_ct.im));
    R a = pow(c_abs(r_ort), (R)2.0f);
    R b = pow(c_abs(r_par), (R)2);
    return (R)0.5f * (a + b);
}

/* fresnel_MyComplex_R (ORF/Fresnel.h:78-91): unpolarised reflectance of a
 * conductor with relative complex IOR eta (= n1 / ior). */
v3 orc_fresnel_conductor(v3 w_i, v3 w_m, const cplx3 *eta)
{
    R cos_i = vdot(w_i, w_m);
    R sin_i = sqrt((R)1 - cos_i * cos_i);
    return V3(fresnel_conductor_1(cos_i, sin_i, eta->x),
              fresnel_conductor_1(cos_i, sin_i, eta->y),
              fresnel_conductor_1(cos_i, sin_i, eta->z));
}

/* fresnel_R(cos, ior1/ior2) (ORF/Fresnel.h:14-40): dielectric, unpolarised. */
R orc_fresnel_dielectric(R cos_theta, R eta)
{
    R sin_t2 = eta * eta * ((R)1 - cos_theta * cos_theta);
    if (sin_t2 >= (R)1)
        return 1;
    R cos_t = sqrt((R)1 - sin_t2);
    R rs = (eta * cos_theta - cos_t) / (eta * cos_theta + cos_t);
    R rp = (cos_theta - eta * cos_t) / (cos_theta + eta * cos_t);
    return (rs * rs + rp * rp) * (R)0.5f;
}

/* ------------------------------------------------------------ operator */

/* Rough-conductor sampling step of CU/metallic_shader.cu:84-158 (visible
 * normals, GGX, isotropic alpha), for a surface with normal (0,0,1).  wo is the
 * reference's w_i (towards the viewer); *wi is its w_o (the reflected ray).
 * weight = F(wo.m) G1(wi) (:117,152); pdf = D_wo(m) / (4 |wi.m|) through
 * microfacet_eval_visible_normal (ORF/Microfacet.h:85-97). */
int orc_ggx_conductor_query(v3 wo, R alpha, R u1, R u2, const cplx3 *eta,
                            v3 *wi, R *pdf, v3 *weight)
{
    const v3 normal = V3(0, 0, 1);
    v3 w_i = wo;
    v3 ffn = orc_faceforward(normal, w_i, normal);
    R ax = alpha, ay = alpha;

    v3 m = orc_to_world(orc_ggx_sample_vndf(orc_to_local(w_i, ffn), ax, ay, u1, u2), ffn);
    R G_i = orc_ggx_G1(w_i, m, ffn, ax, ay);
    v3 F = orc_fresnel_conductor(w_i, m, eta);
    v3 w_o = vadd(vneg(w_i), vscale(vscale(m, (R)2.0f), vdot(w_i, m)));
    R G_o = orc_ggx_G1(w_o, m, ffn, ax, ay);

    R D = orc_ggx_eval_vndf(orc_to_local(w_i, ffn), orc_to_local(m, ffn), ax, ay);
    R dm = vdot(w_o, m);
    *wi = w_o;
    *pdf = D / ((R)4 * (dm < 0 ? -dm : dm));
    if (G_i <= 0 || G_o <= 0) {
        *weight = V3(0, 0, 0);
        return 0;
    }
    *weight = vscale(F, G_o);
    return 1;
}

/* ------------------------------------------------------------ erfinv (FP64) */

/* erfinv (ORF/helpers.h:301-377): Giles (2010), "Approximating the erfinv
 * function", double-precision branches w < 6.25, w < 16, else. */
double orc_erfinv(double x)
{
    double w = -log((1.0 - x) * (1.0 + x)), p;
    if (w < 6.25) {
        w = w - 3.125;
        p = -3.6444120640178196996e-21;
        p = -1.685059138182016589e-19 + p * w;
        p = 1.2858480715256400167e-18 + p * w;
        p = 1.115787767802518096e-17 + p * w;
        p = -1.333171662854620906e-16 + p * w;
        p = 2.0972767875968561637e-17 + p * w;
        p = 6.6376381343583238325e-15 + p * w;
        p = -4.0545662729752068639e-14 + p * w;
        p = -8.1519341976054721522e-14 + p * w;
        p = 2.6335093153082322977e-12 + p * w;
        p = -1.2975133253453532498e-11 + p * w;
        p = -5.4154120542946279317e-11 + p * w;
        p = 1.051212273321532285e-09 + p * w;
        p = -4.1126339803469836976e-09 + p * w;
        p = -2.9070369957882005086e-08 + p * w;
        p = 4.2347877827932403518e-07 + p * w;
        p = -1.3654692000834678645e-06 + p * w;
        p = -1.3882523362786468719e-05 + p * w;
        p = 0.0001867342080340571352 + p * w;
        p = -0.00074070253416626697512 + p * w;
        p = -0.0060336708714301490533 + p * w;
        p = 0.24015818242558961693 + p * w;
        p = 1.6536545626831027356 + p * w;
    } else if (w < 16.0) {
        w = sqrt(w) - 3.25;
        p = 2.2137376921775787049e-09;
        p = 9.0756561938885390979e-08 + p * w;
        p = -2.7517406297064545428e-07 + p * w;
        p = 1.8239629214389227755e-08 + p * w;
        p = 1.5027403968909827627e-06 + p * w;
        p = -4.013867526981545969e-06 + p * w;
        p = 2.9234449089955446044e-06 + p * w;
        p = 1.2475304481671778723e-05 + p * w;
        p = -4.7318229009055733981e-05 + p * w;
        p = 6.8284851459573175448e-05 + p * w;
        p = 2.4031110387097893999e-05 + p * w;
        p = -0.0003550375203628474796 + p * w;
        p = 0.00095328937973738049703 + p * w;
        p = -0.0016882755560235047313 + p * w;
        p = 0.0024914420961078508066 + p * w;
        p = -0.0037512085075692412107 + p * w;
        p = 0.005370914553590063617 + p * w;
        p = 1.0052589676941592334 + p * w;
        p = 3.0838856104922207635 + p * w;
    } else {
        w = sqrt(w) - 5.0;
        p = -2.7109920616438573243e-11;
        p = -2.5556418169965252055e-10 + p * w;
        p = 1.5076572693500548083e-09 + p * w;
        p = -3.7894654401267369937e-09 + p * w;
        p = 7.6157012080783393804e-09 + p * w;
        p = -1.4960026627149240478e-08 + p * w;
        p = 2.9147953450901080826e-08 + p * w;
        p = -6.7711997758452339498e-08 + p * w;
        p = 2.2900482228026654717e-07 + p * w;
        p = -9.9298272942317002539e-07 + p * w;
        p = 4.5260625972231537039e-06 + p * w;
        p = -1.9681778105531670567e-05 + p * w;
        p = 7.5995277030017761139e-05 + p * w;
        p = -0.00021503011930044477347 + p * w;
        p = -0.00013871931833623122026 + p * w;
        p = 1.0103004648645343977 + p * w;
        p = 4.8499064014085844221 + p * w;
    }
    return p * x;
}

/* ------------------------------------------------------------ Beckmann */

#define M_2_SQRTPI_R ((R)1.12837916709551257390f)

/* beckmann_alpha_i (ORF/MicrofacetBeckmann.h:13-21) equals ggx_alpha_i. */
R orc_beckmann_alpha_i(v3 wi, R ax, R ay) { return orc_ggx_alpha_i(wi, ax, ay); }

/* beckmann_G1(cos, alpha^2) (ORF/MicrofacetBeckmann.h:68-75): Walter et al. 2007
 * rational fit of the Smith masking; the denominator's 2.577 is a double literal. */
R orc_beckmann_G1_cos(R cos_theta, R alpha_sqr)
{
    R c2 = cos_theta * cos_theta;
    R t2 = ((R)1 - c2) / c2;
    R a_sqr = (R)1 / (alpha_sqr * t2);
    R a = sqrt(a_sqr);
    if (!(a < (R)1.6f))
        return 1;
    R num = (R)3.535f * a + (R)2.181f * a_sqr;
    double den = (double)((R)1.0f + (R)2.276f * a) + 2.577 * (double)a_sqr;
    return (R)((double)num / den);
}

/* beckmann_G1(v, m, n, ax, ay) (ORF/MicrofacetBeckmann.h:89-98). */
R orc_beckmann_G1(v3 v, v3 m, v3 n, R ax, R ay)
{
    R width = orc_beckmann_alpha_i(orc_to_local(v, n), ax, ay);
    R cos_v = vdot(v, n);
    if (vdot(v, m) * cos_v <= 0)
        return 0;
    return orc_beckmann_G1_cos(cos_v, width * width);
}

/* beckmann_projected_area (ORF/MicrofacetBeckmann.h:45-58). */
R orc_beckmann_projected_area(v3 wi, R ax, R ay)
{
    if (wi.z > (R)0.9999f)
        return 1;
    if (wi.z < (R)-0.9999f)
        return 0;
    R alpha = orc_beckmann_alpha_i(wi, ax, ay);
    R theta = F_ACOS(wi.z);
    R a = (R)1 / F_TAN(theta) / alpha;
    double t1 = (double)((R)0.5f * ((R)F_ERF(a) + (R)1.0f) * wi.z);
    double t2 = 0.25 * (double)M_2_SQRTPI_R * (double)alpha * (double)F_SIN(theta) * (double)F_EXP(-a * a);
    return (R)(t1 + t2);
}

/* beckmann_slope_pdf + beckmann_eval_NDF + beckmann_eval_VNDF (ORF/MicrofacetBeckmann.h:61-65, 249-281). */
R orc_beckmann_eval_ndf(v3 wm, R ax, R ay)
{
    if (wm.z <= 0)
        return 0;
    R sx = -wm.x / wm.z;
    R sy = -wm.y / wm.z;
    R p22 = (R)1 / (PI_R * ax * ay) * F_EXP(-sx * sx / (ax * ax) - sy * sy / (ay * ay));
    return p22 / (wm.z * wm.z * wm.z * wm.z);
}

R orc_beckmann_eval_vndf(v3 wi, v3 wm, R ax, R ay)
{
    if (wm.z <= 0)
        return 0;
    R area = orc_beckmann_projected_area(wi, ax, ay);
    if (area == 0)
        return 0;
    R c = (R)1 / area;
    R D = orc_beckmann_eval_ndf(wm, ax, ay);
    R d = vdot(wi, wm);
    return c * (d > 0 ? d : 0) * D;
}

/* beckmann_sample_P22_11 (ORF/MicrofacetBeckmann.h:161-210): Newton-bisection
 * inversion of the visible-slope CDF in erf space (Heitz & d'Eon 2014).  The
 * reference's loop has no iteration bound; `max_iter` (256) only guards
 * against a non-terminating input and is never reached on finite inputs. */
v2 orc_beckmann_sample_p22_11(R theta_i, R z1, R z2)
{
    v2 slope;
    if (theta_i < (R)0.0001f) {
        R r = sqrt(-F_LOG(z1));
        R phi = PI_R * 2 * z2;
        slope.x = r * F_COS(phi);
        slope.y = r * F_SIN(phi);
        return slope;
    }
    R sin_i = F_SIN(theta_i);
    R cos_i = F_COS(theta_i);
    R slope_i = cos_i / sin_i;
    R a = cos_i / sin_i;
    R proj_area = (R)0.5f * ((R)F_ERF(a) + (R)1.0f) * cos_i + (R)0.25f * M_2_SQRTPI_R * sin_i * F_EXP(-a * a);
    if (proj_area < (R)0.0001f || proj_area != proj_area) {
        slope.x = slope.y = 0;
        return slope;
    }
    R c = (R)1 / proj_area;
    R erf_min = (R)-0.9999f;
    R e = (R)F_ERF(slope_i);
    R erf_max = erf_min < e ? e : erf_min; /* max(erf_min, erf(slope_i)) */
    R erf_cur = (R)0.5f * (erf_min + erf_max);
    for (int it = 0; it < 256 && erf_max - erf_min > (R)0.00001f; it++) {
        if (!(erf_cur >= erf_min && erf_cur <= erf_max))
            erf_cur = (R)0.5f * (erf_min + erf_max);
        R s = (R)orc_erfinv((double)erf_cur);
        R cdf;
        if (s >= slope_i) {
            cdf = 1;
        } else {
            double A = 0.25 * (double)M_2_SQRTPI_R * (double)sin_i * (double)F_EXP(-s * s);
            R B = cos_i * ((R)0.5f + (R)0.5f * (R)F_ERF(s));
            cdf = (R)((double)c * (A + (double)B));
        }
        R diff = cdf - z1;
        if (fabs(diff) < (R)0.00001f)
            break;
        if (diff > 0) {
            if (erf_max == erf_cur)
                break;
            erf_max = erf_cur;
        } else {
            if (erf_min == erf_cur)
                break;
            erf_min = erf_cur;
        }
        R deriv = (R)0.5f * c * cos_i - (R)0.5f * c * sin_i * s;
        erf_cur -= diff / deriv;
    }
    R lo = erf_min < erf_cur ? erf_cur : erf_min;     /* max(erf_min, erf_cur) */
    R cl = lo < erf_max ? lo : erf_max;               /* min(erf_max, .) */
    slope.x = (R)orc_erfinv((double)cl);
    slope.y = (R)orc_erfinv((double)((R)2.0f * z2 - (R)1.0f));
    return slope;
}

/* beckmann_sample_VNDF (ORF/MicrofacetBeckmann.h:213-231). */
v3 orc_beckmann_sample_vndf(v3 wi, R ax, R ay, R z1, R z2)
{
    v3 wi11 = orc_normalize(V3(ax * wi.x, ay * wi.y, wi.z));
    v2 s11 = orc_beckmann_sample_p22_11(F_ACOS(wi11.z), z1, z2);
    R phi = F_ATAN2(wi11.y, wi11.x);
    R cp = F_COS(phi), sp = F_SIN(phi);
    R sx = cp * s11.x - sp * s11.y;
    R sy = sp * s11.x + cp * s11.y;
    sx *= ax;
    sy *= ay;
    if (!isfinite(sx) || sx != sx) {
        if (wi.z > 0)
            return V3(0, 0, 1);
        return orc_normalize(V3(wi.x, wi.y, 0));
    }
    return orc_normalize(V3(-sx, -sy, (R)1));
}

/* fresnel_R(cos1, cos2, ior1/ior2) (ORF/Fresnel.h:26-32). */
R orc_fresnel_R3(R cos1, R cos2, R eta)
{
    R rs = (eta * cos1 - cos2) / (eta * cos1 + cos2);
    R rp = (cos1 - eta * cos2) / (cos1 + eta * cos2);
    return (rs * rs + rp * rp) * (R)0.5f;
}

/* ---- microsurface multiple scattering: Heitz et al. 2016 random walk on a
 * Gaussian-height GGX surface (ORF/MicrofacetGGX.h:19-29, 93-107, 254-305,
 * 412-438; heights ORF/helpers.h:380-391). */

#define RAY_MAX_R ((R)1e27f)

/* ggx_smith_lambda (ORF/MicrofacetGGX.h:20-29). */
R orc_ggx_smith_lambda(v3 wi, R ax, R ay)
{
    if (wi.z > (R)0.99999f) return 0;
    if (wi.z < (R)-0.99999f) return -1;
    R theta = F_ACOS(wi.z);
    R a = (R)1.0f / F_TAN(theta) / orc_ggx_alpha_i(wi, ax, ay);
    return (R)0.5f * ((R)-1.0f + sqrt((R)1 + (R)1 / (a * a)) * (a > 0 ? (R)1 : (R)-1));
}

/* GaussianHeightCDF / GaussianHeightInvCDF (ORF/helpers.h:380-391): erf in R,
 * erfinv in double rounded to R, as the reference mixes them. */
R orc_height_cdf(R h) { return (R)0.5f + (R)0.5f * (R)F_ERF((R)0.707106781186547524401f * h); }
R orc_height_inv_cdf(R z) { return (R)1.41421356237309504880f * (R)orc_erfinv((double)((R)2.0f * z - (R)1.0f)); }

/* ggx_G(wi, h) (ORF/MicrofacetGGX.h:93-107). */
R orc_ggx_masking_height(v3 wi, R h, R ax, R ay)
{
    if (wi.z > (R)0.9999f) return 1;
    if (wi.z <= 0) return 0;
    return F_POW(orc_height_cdf(h), orc_ggx_smith_lambda(wi, ax, ay));
}

/* ggx_sample_height (ORF/MicrofacetGGX.h:254-275). */
R orc_ggx_sample_height(v3 wr, R hr, R z, R ax, R ay)
{
    if (wr.z > (R)0.9999f) return RAY_MAX_R;
    if (wr.z < (R)-0.9999f) return orc_height_inv_cdf(z * orc_height_cdf(hr));
    if (fabs(wr.z) < (R)0.00001f) return hr;
    R g1 = orc_ggx_masking_height(wr, hr, ax, ay);
    if (z > (R)1.0f - g1) return RAY_MAX_R;
    return orc_height_inv_cdf(orc_height_cdf(hr) / F_POW((R)1.0f - z, (R)1.0f / orc_ggx_smith_lambda(wr, ax, ay)));
}

/* ggx_sample_dielectric_phase (ORF/MicrofacetGGX.h:278-305); *outside flips on
 * refraction (the reference passes one flag as both wi_outside and wo_outside). */
static v3 ggx_sample_dielectric_phase(v3 w_i, R m_eta, uint32_t* seed, int* outside, R ax, R ay)
{
    R z1 = orc_rnd_tea(seed);
    R z2 = orc_rnd_tea(seed);
    int wi_out = *outside;
    R eta = wi_out ? m_eta : (R)1.0f / m_eta;
    v3 w_m = wi_out ? orc_ggx_sample_vndf(w_i, ax, ay, z1, z2) : vneg(orc_ggx_sample_vndf(vneg(w_i), ax, ay, z1, z2));
    R c = fabs(vdot(w_m, w_i));
    R st2 = eta * eta * ((R)1.0f - c * c);
    R ct = sqrt((R)1.0f - st2);
    R F = st2 < (R)1.0f ? orc_fresnel_R3(c, ct, eta) : (R)1.0f;
    if (orc_rnd_tea(seed) <= F) return orc_normalize(vadd(vneg(w_i), vscale(vscale(w_m, (R)2.0f), vdot(w_i, w_m))));
    *outside = !wi_out;
    return orc_normalize(vadd(vscale(w_i, -eta), vscale(w_m, eta * c - ct)));
}

/* beckmann_smith_lambda (ORF/MicrofacetBeckmann.h:24-38). */
R orc_beckmann_smith_lambda(v3 wi, R ax, R ay)
{
    if (wi.z > (R)0.9999f) return 0;
    if (wi.z < (R)-0.9999f) return -1;
    R th = F_ACOS(wi.z);
    R t = F_TAN(th);
    if (isinf(t)) return 0;
    R a = (R)1.0f / (t * orc_ggx_alpha_i(wi, ax, ay));
    return (R)0.5f * ((R)F_ERF(a) - (R)1.0f) + (R)0.25f * M_2_SQRTPI_R / a * F_EXP(-a * a);
}

/* beckmann_sample_height (ORF/MicrofacetBeckmann.h:284-306): 0.999 direction
 * limits; escape probability 1 / (1 + Lambda) (beckmann_G(wi, a_x, a_y), :130-141). */
R orc_beckmann_sample_height(v3 wr, R hr, R z, R ax, R ay)
{
    if (wr.z > (R)0.999f) return RAY_MAX_R;
    if (wr.z < (R)-0.999f) return orc_height_inv_cdf(z * orc_height_cdf(hr));
    if (fabs(wr.z) < (R)0.00001f) return hr;
    R g1 = wr.z > (R)0.9999f ? (R)1 : (wr.z <= 0 ? (R)0 : (R)1.0f / ((R)1.0f + orc_beckmann_smith_lambda(wr, ax, ay)));
    if (z > (R)1.0f - g1) return RAY_MAX_R;
    R p = F_POW((R)1.0f - z, (R)1.0f / orc_beckmann_smith_lambda(wr, ax, ay));
    return orc_height_inv_cdf(orc_height_cdf(hr) / p);
}

/* beckmann_sample_dielectric_phase (ORF/MicrofacetBeckmann.h:310-340). */
static v3 beckmann_sample_dielectric_phase(v3 w_i, R m_eta, uint32_t* seed, int* outside, R ax, R ay)
{
    R z1 = orc_rnd_tea(seed);
    R z2 = orc_rnd_tea(seed);
    int wi_out = *outside;
    R eta = wi_out ? m_eta : (R)1.0f / m_eta;
    v3 wm = wi_out ? orc_beckmann_sample_vndf(w_i, ax, ay, z1, z2) : vneg(orc_beckmann_sample_vndf(vneg(w_i), ax, ay, z1, z2));
    R c = fabs(vdot(wm, w_i));
    R st2 = eta * eta * ((R)1.0f - c * c);
    R ct = sqrt((R)1.0f - st2);
    R F = st2 < (R)1.0f ? orc_fresnel_R3(c, ct, eta) : (R)1.0f;
    if (orc_rnd_tea(seed) <= F) return orc_normalize(vadd(vneg(w_i), vscale(vscale(wm, (R)2.0f), vdot(w_i, wm))));
    *outside = !wi_out;
    return orc_normalize(vadd(vscale(w_i, -eta), vscale(wm, eta * c - ct)));
}

/* beckmann_multiscattering_dielectric_BSDF_sample (ORF/MicrofacetBeckmann.h:445-475). */
int orc_beckmann_ms_dielectric_sample(v3 local_wi, R eta, R ax, R ay, uint32_t* seed, int cap, v3* w_out, int* order)
{
    R hr = (R)1.0f + orc_height_inv_cdf((R)0.9999f);
    v3 w_o = vneg(local_wi);
    int outside = (-w_o.z) > 0;
    *order = 0;
    for (;;) {
        R z = orc_rnd_tea(seed);
        hr = outside ? orc_beckmann_sample_height(w_o, hr, z, ax, ay)
                     : -orc_beckmann_sample_height(vneg(w_o), -hr, z, ax, ay);
        if (hr == RAY_MAX_R || hr == -RAY_MAX_R) break;
        if (++*order > cap) return 0;
        w_o = beckmann_sample_dielectric_phase(vneg(w_o), eta, seed, &outside, ax, ay);
        if (hr != hr || w_o.z != w_o.z) {
            w_o = V3(0, 0, 1);
            break;
        }
    }
    *w_out = w_o;
    return 1;
}

/* beckmann_multiscattering_conductor_BSDF_sample (ORF/MicrofacetBeckmann.h:711-736). */
int orc_beckmann_ms_conductor_sample(v3 local_wi, const cplx3* eta, const cplx3* ieta, R ax, R ay, uint32_t* seed,
                                     int cap, v3* w_out, v3* weight, int* order)
{
    R hr = (R)1.0f + orc_height_inv_cdf((R)0.9999f);
    v3 w_o = vneg(local_wi);
    *weight = V3(1, 1, 1);
    *order = 0;
    for (;;) {
        R u = orc_rnd_tea(seed);
        hr = orc_beckmann_sample_height(w_o, hr, u, ax, ay);
        if (hr == RAY_MAX_R) break;
        if (++*order > cap) return 0;
        v3 w_t = vneg(w_o);
        R z1 = orc_rnd_tea(seed);
        R z2 = orc_rnd_tea(seed);
        v3 wm = orc_beckmann_sample_vndf(w_t, ax, ay, z1, z2);
        w_o = vadd(vneg(w_t), vscale(vscale(wm, (R)2.0f), vdot(w_t, wm)));
        *weight = vmul(*weight, orc_fresnel_conductor(w_t, wm, w_t.z > 0 ? eta : ieta));
        if (hr != hr || w_o.z != w_o.z) {
            w_o = V3(0, 0, 1);
            break;
        }
    }
    *w_out = w_o;
    return 1;
}

/* ggx_multiscattering_conductor_BSDF_sample (ORF/MicrofacetGGX.h:614-637): the
 * random walk on the GGX microsurface of a conductor, local frame.  Each event
 * reflects about a visible normal of the arriving direction w_t and multiplies
 * the weight by the complex Fresnel reflectance with eta (w_t.z > 0) or 1/eta.
 * Returns 0 when the walk exceeds `cap` events (new: the reference's loop has no
 * bound). */
int orc_ggx_ms_conductor_sample(v3 local_wi, const cplx3* eta, const cplx3* ieta, R ax, R ay, uint32_t* seed, int cap,
                                v3* w_out, v3* weight, int* order)
{
    R hr = (R)1.0f + orc_height_inv_cdf((R)0.9999f);
    v3 w_o = vneg(local_wi);
    *weight = V3(1, 1, 1);
    *order = 0;
    for (;;) {
        R u = orc_rnd_tea(seed);
        hr = orc_ggx_sample_height(w_o, hr, u, ax, ay);
        if (hr == RAY_MAX_R) break;
        if (++*order > cap) return 0;
        v3 w_t = vneg(w_o);
        R z1 = orc_rnd_tea(seed);
        R z2 = orc_rnd_tea(seed);
        v3 wm = orc_ggx_sample_vndf(w_t, ax, ay, z1, z2);
        w_o = vadd(vneg(w_t), vscale(vscale(wm, (R)2.0f), vdot(w_t, wm)));
        *weight = vmul(*weight, orc_fresnel_conductor(w_t, wm, w_t.z > 0 ? eta : ieta));
        if (hr != hr || w_o.z != w_o.z) {
            w_o = V3(0, 0, 1);
            break;
        }
    }
    *w_out = w_o;
    return 1;
}

/* ggx_multiscattering_dielectric_BSDF_sample (ORF/MicrofacetGGX.h:412-438) with
 * at most `cap` scattering events (the reference's loop is unbounded); returns
 * 0 when the walk exceeds the cap.  *order receives the event count. */
int orc_ggx_ms_dielectric_sample(v3 local_wi, R eta, R ax, R ay, uint32_t* seed, int cap, v3* w_out, int* order)
{
    R hr = (R)1.0f + orc_height_inv_cdf((R)0.9999f);
    v3 w_o = vneg(local_wi);
    int outside = (-w_o.z) > 0;
    *order = 0;
    for (;;) {
        R z = orc_rnd_tea(seed);
        hr = outside ? orc_ggx_sample_height(w_o, hr, z, ax, ay) : -orc_ggx_sample_height(vneg(w_o), -hr, z, ax, ay);
        if (hr == RAY_MAX_R || hr == -RAY_MAX_R) break;
        if (++*order > cap) return 0;
        w_o = ggx_sample_dielectric_phase(vneg(w_o), eta, seed, &outside, ax, ay);
        if (hr != hr || w_o.z != w_o.z) {
            w_o = V3(0, 0, 1);
            break;
        }
    }
    *w_out = w_o;
    return 1;
}

/* fresnel_R(w_i, w_m, ior1/ior2) (ORF/Fresnel.h:42-53). */
R orc_fresnel_R_dir(v3 w_i, v3 w_m, R eta)
{
    R c1 = vdot(w_i, w_m);
    R c22 = (R)1.0f - eta * eta * ((R)1.0f - c1 * c1);
    if (c22 <= 0) return 1;
    return orc_fresnel_R3(c1, sqrt(c22), eta);
}

/* ggx_eval_dielectric_phase (ORF/MicrofacetGGX.h:308-339). */
R orc_ggx_eval_dielectric_phase(v3 w_i, v3 w_o, int wi_out, int wo_out, R m_eta, R ax, R ay)
{
    R eta = wi_out ? m_eta : (R)1.0f / m_eta;
    if (wi_out == wo_out) {
        v3 w_h = orc_normalize(vadd(w_i, w_o));
        R D = wi_out ? orc_ggx_eval_vndf(w_i, w_h, ax, ay) : orc_ggx_eval_vndf(vneg(w_i), vneg(w_h), ax, ay);
        R F = wi_out ? orc_fresnel_R_dir(w_i, w_h, eta) : orc_fresnel_R_dir(vneg(w_i), vneg(w_h), eta);
        R a = wi_out ? (R)0.25f / vdot(w_i, w_h) : (R)0.25f / vdot(vneg(w_i), vneg(w_h));
        return F * D * a;
    }
    v3 w_h = vneg(orc_normalize(vadd(w_i, V3(w_o.x / eta, w_o.y / eta, w_o.z / eta))));
    w_h = vscale(w_h, wi_out ? w_h.z / fabs(w_h.z) : -w_h.z / fabs(w_h.z));
    if (vdot(w_h, w_i) < 0) return 0;
    R D = wi_out ? orc_ggx_eval_vndf(w_i, w_h, ax, ay) : orc_ggx_eval_vndf(vneg(w_i), vneg(w_h), ax, ay);
    R F = wi_out ? orc_fresnel_R_dir(w_i, w_h, eta) : orc_fresnel_R_dir(vneg(w_i), vneg(w_h), eta);
    R b = (R)1.0f / (eta * eta);
    R cc = wi_out ? -vdot(w_o, w_h) : -vdot(vneg(w_o), vneg(w_h));
    R c = cc > 0 ? cc : 0;
    R e = wi_out ? vdot(w_i, w_h) + vdot(w_o, w_h) / eta : vdot(vneg(w_i), vneg(w_h)) + vdot(vneg(w_o), vneg(w_h)) / eta;
    R d = e * e;
    R a = b * c / d;
    return a * D * ((R)1.0f - F);
}

/* ggx_multiscattering_dielectric_BSDF_eval (ORF/MicrofacetGGX.h:444-478): one
 * random-walk estimate of BSDF x cos, summed over the walk's events; truncated
 * after `cap` events (the orders the sampler keeps; the reference's loop has
 * no bound). */
R orc_ggx_ms_dielectric_eval(v3 w_i, v3 w_o, R eta, R ax, R ay, uint32_t* seed, int cap)
{
    v3 w_r = vneg(w_i);
    R hr = (R)1.0f + orc_height_inv_cdf((R)0.9999f);
    int order = 0;
    R sum = 0;
    int wr_out = (-w_r.z) > 0;
    int wo_out = w_o.z > 0;
    for (;;) {
        R z = orc_rnd_tea(seed);
        hr = wr_out ? orc_ggx_sample_height(w_r, hr, z, ax, ay) : -orc_ggx_sample_height(vneg(w_r), -hr, z, ax, ay);
        if (hr == RAY_MAX_R || hr == -RAY_MAX_R || isinf(hr)) break;
        order++;
        R phase = orc_ggx_eval_dielectric_phase(vneg(w_r), w_o, wr_out, wo_out, eta, ax, ay);
        R shadowing = wo_out ? orc_ggx_masking_height(w_o, hr, ax, ay) : orc_ggx_masking_height(vneg(w_o), -hr, ax, ay);
        sum += phase * shadowing;
        w_r = ggx_sample_dielectric_phase(vneg(w_r), eta, seed, &wr_out, ax, ay);
        if (hr != hr || w_r.z != w_r.z) {
            sum = 0;
            break;
        }
        if (order >= cap) break;
    }
    return sum;
}

/* ---- the walks of ORF/Microfacet.h:101-184 for either distribution (dist 1:
 * GGX, 0: Beckmann). */

static R ms_sample_height(int dist, v3 wr, R hr, R z, R ax, R ay)
{
    return dist == 1 ? orc_ggx_sample_height(wr, hr, z, ax, ay) : orc_beckmann_sample_height(wr, hr, z, ax, ay);
}

/* ggx_G(wi, h) / beckmann_G(wi, h) (ORF/MicrofacetGGX.h:93-107,
 * ORF/MicrofacetBeckmann.h:108-127). */
R orc_ms_masking_height(int dist, v3 wi, R h, R ax, R ay)
{
    if (wi.z > (R)0.9999f) return 1;
    if (wi.z <= 0) return 0;
    R lambda = dist == 1 ? orc_ggx_smith_lambda(wi, ax, ay) : orc_beckmann_smith_lambda(wi, ax, ay);
    return F_POW(orc_height_cdf(h), lambda);
}

static v3 ms_sample_vndf(int dist, v3 wi, R ax, R ay, R z1, R z2)
{
    return dist == 1 ? orc_ggx_sample_vndf(wi, ax, ay, z1, z2) : orc_beckmann_sample_vndf(wi, ax, ay, z1, z2);
}

/* ggx_sample_diffuse_phase / beckmann_sample_diffuse_phase (ORF/MicrofacetGGX.h:
 * 342-375, ORF/MicrofacetBeckmann.h:375-408): visible normal (z1, z2), then the
 * concentric disk map (Shirley and Chiu 1997) of (z3, z4) lifted to the
 * hemisphere about it in its create_onb frame. */
static v3 ms_sample_diffuse_phase(int dist, v3 w_i, R ax, R ay, uint32_t* seed, v3* w_m)
{
    R z1 = orc_rnd_tea(seed);
    R z2 = orc_rnd_tea(seed);
    R z3 = orc_rnd_tea(seed);
    R z4 = orc_rnd_tea(seed);
    *w_m = ms_sample_vndf(dist, w_i, ax, ay, z1, z2);
    v3 w1, w2;
    orc_create_onb(*w_m, &w1, &w2);
    R r1 = (R)2.0f * z3 - (R)1.0f;
    R r2 = (R)2.0f * z4 - (R)1.0f;
    R r, phi;
    if (r1 == 0 && r2 == 0) {
        r = phi = 0;
    } else if (r1 * r1 > r2 * r2) {
        r = r1;
        phi = (PI_R / (R)4.0f) * (r2 / r1);
    } else {
        r = r2;
        phi = (PI_R / (R)2.0f) - (r1 / r2) * (PI_R / (R)4.0f);
    }
    R x = r * F_COS(phi);
    R y = r * F_SIN(phi);
    R zz = (R)1.0f - x * x - y * y;
    R z = sqrt(zz > 0 ? zz : (R)0);
    return vadd(vadd(vscale(w1, x), vscale(w2, y)), vscale(*w_m, z));
}

/* ggx_multiscattering_diffuse_BSDF_sample / beckmann_... (ORF/MicrofacetGGX.h:
 * 519-544, ORF/MicrofacetBeckmann.h:613-640): the walk on a Lambertian
 * microsurface, weight rho per event, at most `cap` events (the reference
 * stops at 100 and returns the last sampled direction as it is); *w_m is the
 * last event's visible normal. */
void orc_ms_diffuse_sample(int dist, v3 local_wi, R ax, R ay, v3 rho, uint32_t* seed, int cap, v3* w_out, v3* w_m,
                           v3* weight, int* order)
{
    *weight = V3(1, 1, 1);
    *w_m = V3(0, 0, 1);
    *order = 0;
    R hr = (R)1.0f + orc_height_inv_cdf((R)0.9999f);
    v3 w_o = vneg(local_wi);
    while (*order < cap) {
        R u = orc_rnd_tea(seed);
        hr = ms_sample_height(dist, w_o, hr, u, ax, ay);
        if (hr == RAY_MAX_R) break;
        ++*order;
        *weight = vmul(*weight, rho);
        w_o = ms_sample_diffuse_phase(dist, vneg(w_o), ax, ay, seed, w_m);
        if (hr != hr || w_o.z != w_o.z) {
            w_o = V3(0, 0, 1);
            break;
        }
    }
    *w_out = w_o;
}

/* ggx_multiscattering_diffuse_BSDF_eval / beckmann_... (ORF/MicrofacetGGX.h:
 * 547-577, ORF/MicrofacetBeckmann.h:643-674): sum over the walk's events of
 * rho (once per event, as written there) x max(0, w_o.m) / pi for a visible
 * normal m drawn afresh x the masking of w_o at the event's height, after at
 * most `cap` events. */
v3 orc_ms_diffuse_eval(int dist, v3 w_i, v3 w_o, R ax, R ay, v3 rho, uint32_t* seed, int cap)
{
    v3 w_r = vneg(w_i);
    R hr = (R)1.0f + orc_height_inv_cdf((R)0.9999f);
    v3 sum = V3(0, 0, 0);
    for (int order = 0; order < cap; ++order) {
        R u = orc_rnd_tea(seed);
        hr = ms_sample_height(dist, w_r, hr, u, ax, ay);
        if (hr == RAY_MAX_R) break;
        R z1 = orc_rnd_tea(seed);
        R z2 = orc_rnd_tea(seed);
        v3 m = ms_sample_vndf(dist, vneg(w_r), ax, ay, z1, z2);
        R c = vdot(w_o, m);
        R phase = (R)0.318309886183790671538 * (c > 0 ? c : (R)0); /* M_1_PI */
        R g = orc_ms_masking_height(dist, w_o, hr, ax, ay);
        v3 term = V3(rho.x * (phase * g), rho.y * (phase * g), rho.z * (phase * g));
        if (isfinite(term.x) && isfinite(term.y) && isfinite(term.z)) sum = vadd(sum, term);
        v3 ms;
        w_r = ms_sample_diffuse_phase(dist, vneg(w_r), ax, ay, seed, &ms);
        if (hr != hr || w_r.z != w_r.z) return V3(0, 0, 0);
    }
    return sum;
}

/* ggx_multiscattering_conductor_BSDF_eval / beckmann_... (ORF/MicrofacetGGX.h:
 * 580-611, ORF/MicrofacetBeckmann.h:677-708): sum over the walk's events of
 * F(-w_r, h) x D_{-w_r}(h) / 4 (h the half vector of -w_r and w_o; 0 when h
 * points below the mean plane) x the masking of w_o at the event's height,
 * with eta or ieta as -w_r points up or down; at most `cap` events. */
v3 orc_ms_conductor_eval(int dist, v3 w_i, v3 w_o, const cplx3* eta, const cplx3* ieta, R ax, R ay, uint32_t* seed,
                         int cap)
{
    v3 w_r = vneg(w_i);
    R hr = (R)1.0f + orc_height_inv_cdf((R)0.9999f);
    v3 sum = V3(0, 0, 0);
    for (int order = 0; order < cap; ++order) {
        R u = orc_rnd_tea(seed);
        hr = ms_sample_height(dist, w_r, hr, u, ax, ay);
        if (hr == RAY_MAX_R) break;
        v3 w_t = vneg(w_r);
        v3 h = orc_normalize(vadd(w_t, w_o));
        R phase = 0;
        if (h.z >= 0) phase = (R)0.25f * (dist == 1 ? orc_ggx_eval_vndf(w_t, h, ax, ay) : orc_beckmann_eval_vndf(w_t, h, ax, ay));
        v3 F = orc_fresnel_conductor(w_t, h, w_t.z > 0 ? eta : ieta);
        R g = orc_ms_masking_height(dist, w_o, hr, ax, ay);
        v3 term = vscale(F, phase * g);
        if (isfinite(term.x) && isfinite(term.y) && isfinite(term.z)) sum = vadd(sum, term);
        R z1 = orc_rnd_tea(seed);
        R z2 = orc_rnd_tea(seed);
        v3 m = ms_sample_vndf(dist, w_t, ax, ay, z1, z2);
        w_r = vadd(vneg(w_t), vscale(vscale(m, (R)2.0f), vdot(w_t, m)));
        if (hr != hr || w_r.z != w_r.z) return V3(0, 0, 0);
    }
    return sum;
}

/* beckmann_eval_dielectric_phase (ORF/MicrofacetBeckmann.h:342-372). */
R orc_beckmann_eval_dielectric_phase(v3 w_i, v3 w_o, int wi_out, int wo_out, R m_eta, R ax, R ay)
{
    R eta = wi_out ? m_eta : (R)1.0f / m_eta;
    if (wi_out == wo_out) {
        v3 w_h = orc_normalize(vadd(w_i, w_o));
        R D = wi_out ? orc_beckmann_eval_vndf(w_i, w_h, ax, ay) : orc_beckmann_eval_vndf(vneg(w_i), vneg(w_h), ax, ay);
        R F = wi_out ? orc_fresnel_R_dir(w_i, w_h, eta) : orc_fresnel_R_dir(vneg(w_i), vneg(w_h), eta);
        R a = wi_out ? (R)0.25f / vdot(w_i, w_h) : (R)0.25f / vdot(vneg(w_i), vneg(w_h));
        return F * D * a;
    }
    v3 w_h = vneg(orc_normalize(vadd(w_i, V3(w_o.x / eta, w_o.y / eta, w_o.z / eta))));
    w_h = vscale(w_h, wi_out ? w_h.z / fabs(w_h.z) : -w_h.z / fabs(w_h.z));
    if (vdot(w_h, w_i) < 0) return 0;
    R D = wi_out ? orc_beckmann_eval_vndf(w_i, w_h, ax, ay) : orc_beckmann_eval_vndf(vneg(w_i), vneg(w_h), ax, ay);
    R F = wi_out ? orc_fresnel_R_dir(w_i, w_h, eta) : orc_fresnel_R_dir(vneg(w_i), vneg(w_h), eta);
    R b = (R)1.0f / (eta * eta);
    R cc = wi_out ? -vdot(w_o, w_h) : -vdot(vneg(w_o), vneg(w_h));
    R c = cc > 0 ? cc : 0;
    R e = wi_out ? vdot(w_i, w_h) + vdot(w_o, w_h) / eta : vdot(vneg(w_i), vneg(w_h)) + vdot(vneg(w_o), vneg(w_h)) / eta;
    return b * c / (e * e) * D * ((R)1.0f - F);
}

/* beckmann_multiscattering_dielectric_BSDF_eval (ORF/MicrofacetBeckmann.h:
 * 479-527) summing every order (order argument 0), non-finite terms skipped,
 * after at most `cap` events. */
R orc_beckmann_ms_dielectric_eval(v3 w_i, v3 w_o, R eta, R ax, R ay, uint32_t* seed, int cap)
{
    v3 w_r = vneg(w_i);
    R hr = (R)1.0f + orc_height_inv_cdf((R)0.9999f);
    R sum = 0;
    int wr_out = (-w_r.z) > 0;
    int wo_out = w_o.z > 0;
    for (int order = 0; order < cap; ++order) {
        R z = orc_rnd_tea(seed);
        hr = wr_out ? orc_beckmann_sample_height(w_r, hr, z, ax, ay)
                    : -orc_beckmann_sample_height(vneg(w_r), -hr, z, ax, ay);
        if (hr == RAY_MAX_R || hr == -RAY_MAX_R || isinf(hr)) break;
        R phase = orc_beckmann_eval_dielectric_phase(vneg(w_r), w_o, wr_out, wo_out, eta, ax, ay);
        R shadowing = wo_out ? orc_ms_masking_height(0, w_o, hr, ax, ay) : orc_ms_masking_height(0, vneg(w_o), -hr, ax, ay);
        R term = phase * shadowing;
        if (isfinite(term)) sum += term;
        w_r = beckmann_sample_dielectric_phase(vneg(w_r), eta, seed, &wr_out, ax, ay);
        if (hr != hr || w_r.z != w_r.z) return 0;
    }
    return sum;
}

/* masking_G1(v, m, n, alpha, distribution) (ORF/Microfacet.h:16-22): ggx_G1(v,
 * m, n, a_g) / beckmann_G1(v, m, n, width) (ORF/MicrofacetGGX.h:77-83,
 * ORF/MicrofacetBeckmann.h:78-86). */
R orc_masking_iso(v3 v, v3 m, v3 n, R a, int dist)
{
    R cv = vdot(v, n);
    if (dist == 1) {
        if (cv * vdot(v, m) <= 0) return 0;
        return orc_ggx_G1_cos(cv, a * a);
    }
    if (vdot(v, m) * cv <= 0) return 0;
    return orc_beckmann_G1_cos(cv, a * a);
}

/* microfacet_sample_normal (ORF/Microfacet.h:50-60) -> ggx_sample_hemisphere /
 * beckmann_sample_hemisphere (ORF/MicrofacetGGX.h:123-139,
 * ORF/MicrofacetBeckmann.h:144-158). */
v3 orc_sample_normal_walter(v3 n, R z1, R z2, R a, int dist)
{
    R phi = (R)2.0f * PI_R * z2;
    R theta;
    if (dist == 1) {
        R tan_theta = a * sqrt(z1) / sqrt((R)1.0f - z1);
        theta = F_ATAN(tan_theta);
    } else {
        R tan_theta_sqr = -a * a * F_LOG((R)1.0f - z1);
        theta = F_ATAN(sqrt(tan_theta_sqr));
    }
    R ct = F_COS(theta), st = F_SIN(theta);
    v3 l = V3(st * F_COS(phi), st * F_SIN(phi), ct);
    v3 u, v;
    orc_create_onb(n, &u, &v);
    return vadd(vadd(vscale(u, l.x), vscale(v, l.y)), vscale(n, l.z));
}

/* microfacet_eval_visible_normal (ORF/Microfacet.h:85-97). */
R orc_eval_visible_normal(v3 wi, v3 wm, v3 n, R ax, R ay, int dist)
{
    v3 li = orc_to_local(wi, n), lm = orc_to_local(wm, n);
    return dist == 1 ? orc_ggx_eval_vndf(li, lm, ax, ay) : orc_beckmann_eval_vndf(li, lm, ax, ay);
}

/* The shared FP32 transcendentals, exported for their accuracy tests:
 * fn 0 sin, 1 cos, 2 tan, 3 asin, 4 acos, 5 atan, 6 atan2(x, y), 7 log, 8 exp, 9 erf. */
void orc_fmath_eval(int fn, int64_t n, const float* x, const float* y, float* out)
{
    for (int64_t i = 0; i < n; i++) {
        float a = x[i], r;
        switch (fn) {
        case 0: r = orf_sinf(a); break;
        case 1: r = orf_cosf(a); break;
        case 2: r = orf_tanf(a); break;
        case 3: r = orf_asinf(a); break;
        case 4: r = orf_acosf(a); break;
        case 5: r = orf_atanf(a); break;
        case 6: r = orf_atan2f(a, y[i]); break;
        case 7: r = orf_logf(a); break;
        case 8: r = orf_expf(a); break;
        default: r = orf_erff(a); break;
        }
        out[i] = r;
    }
}
