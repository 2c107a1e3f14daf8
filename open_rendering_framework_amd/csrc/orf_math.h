/* orf_math.h -- FP32 transcendental functions shared by the device kernels, the
 * host code and the CPU oracle's FP32 build.
 *
 * The device libm (ocml) and glibc round sin/cos/acos/atan2/log/exp a few ulp
 * apart, and the render path amplifies such differences where it is
 * ill-conditioned (grazing angles, slope sampling near the horizon).  One set of
 * functions, built only from IEEE +,-,*,/ and sqrt (correctly rounded on both
 * sides; compiled without contraction), gives the same bits on the GPU and the
 * CPU.  Methods:
 *   sin/cos     : Cephes sinf/cosf (S. L. Moshier, "Methods and Programs for
 *                 Mathematical Functions", 1989): Cody-Waite reduction by pi/4
 *                 in three parts, degree-3 polynomials in r^2;
 *   asin/acos   : Cephes asinf/acosf (half-angle identity above 0.5);
 *   atan/atan2  : Cephes atanf (reduction at tan(pi/8), tan(3pi/8)) plus the
 *                 quadrant rules of C99 F.9.1.4 for finite arguments;
 *                 offsets by pi, pi/2, pi/4 are added as float head + tail;
 *   log/exp     : Cephes logf/expf (frexp / ldexp done on the bits);
 *   pow         : exp(y log x) with pow's special cases;
 *   erf         : x P(x^2) on [0,1] and 1 - exp(-x^2) Q(u) on [1,3.92]
 *                 (coefficients fitted here, relative-error weighted least
 *                 squares of erf(x)/x and erfc(x) exp(x^2)), 1 beyond.
 * Accuracy is checked against double-precision libm over dense sweeps.
 * Domain: finite arguments of moderate size (|x| <= 2^24 for sin/cos); NaN in
 * gives NaN out, and acos/asin of |x| > 1 and log of x < 0 give NaN as libm does.
 *
 * C-compatible (the oracle is C); in HIP translation units the functions are
 * __host__ __device__.
 */
#ifndef ORF_MATH_H
#define ORF_MATH_H

#include <stdint.h>

#if defined(__HIP__) || defined(__HIPCC__)
#define ORF_MATH_FN static inline __host__ __device__
#else
#define ORF_MATH_FN static inline
#endif

typedef union {
    float f;
    uint32_t u;
} orf_fbits_t;

ORF_MATH_FN uint32_t orf_f2u(float f)
{
    orf_fbits_t t;
    t.f = f;
    return t.u;
}

ORF_MATH_FN float orf_u2f(uint32_t u)
{
    orf_fbits_t t;
    t.u = u;
    return t.f;
}

ORF_MATH_FN float orf_fabsf(float x) { return orf_u2f(orf_f2u(x) & 0x7fffffffu); }

ORF_MATH_FN float orf_nanf(void) { return orf_u2f(0x7fc00000u); }

#define ORF_PIF 3.14159265358979323846f
#define ORF_PIO2F 1.57079632679489661923f
#define ORF_PIO4F 0.78539816339744830962f
/* pi, pi/2, pi/4 as float head + float tail (head + tail = value to 4e-15) */
#define ORF_PI_LO (-8.7422776573e-08f)
#define ORF_PIO2_LO (-4.3711388287e-08f)
#define ORF_PIO4_LO (-2.1855694143e-08f)

/* sin and cos of x (Cephes sinf.c / cosf.c). */
ORF_MATH_FN void orf_sincosf(float x, float* s, float* c)
{
    float ax = orf_fabsf(x);
    if (!(ax <= 16777216.0f)) { /* NaN, infinity or out of the supported range */
        *s = *c = orf_nanf();
        return;
    }
    int32_t j = (int32_t)(1.27323954473516f * ax); /* 4 / pi */
    float y = (float)j;
    if (j & 1) {
        j += 1;
        y += 1.0f;
    }
    int q = (int)(j & 7);
    float r = ((ax - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
    float z = r * r;
    float ps = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * r + r;
    float pc = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z -
               0.5f * z + 1.0f;
    float sv, cv;
    if (q == 0) {
        sv = ps;
        cv = pc;
    } else if (q == 2) {
        sv = pc;
        cv = -ps;
    } else if (q == 4) {
        sv = -ps;
        cv = -pc;
    } else {
        sv = -pc;
        cv = ps;
    }
    *s = (orf_f2u(x) >> 31) ? -sv : sv;
    *c = cv;
}

ORF_MATH_FN float orf_sinf(float x)
{
    float s, c;
    orf_sincosf(x, &s, &c);
    return s;
}

ORF_MATH_FN float orf_cosf(float x)
{
    float s, c;
    orf_sincosf(x, &s, &c);
    return c;
}

ORF_MATH_FN float orf_tanf(float x)
{
    float s, c;
    orf_sincosf(x, &s, &c);
    return s / c;
}

/* asin of |x| <= 1 without the sign (Cephes asinf.c). */
ORF_MATH_FN float orf_asin_abs(float a)
{
    float x, z;
    int flag;
    if (a > 0.5f) {
        z = 0.5f * (1.0f - a);
        x = __builtin_sqrtf(z);
        flag = 1;
    } else {
        x = a;
        z = x * x;
        flag = 0;
    }
    z = ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z +
         1.6666752422e-1f) * z * x + x;
    if (flag) z = ORF_PIO2F - ((z + z) - ORF_PIO2_LO);
    return z;
}

ORF_MATH_FN float orf_asinf(float x)
{
    float a = orf_fabsf(x);
    if (!(a <= 1.0f)) return orf_nanf();
    float r = orf_asin_abs(a);
    return x < 0.0f ? -r : r;
}

/* acos (Cephes acosf.c). */
ORF_MATH_FN float orf_acosf(float x)
{
    if (!(orf_fabsf(x) <= 1.0f)) return orf_nanf();
    if (x < -0.5f) return ORF_PIF - (2.0f * orf_asin_abs(__builtin_sqrtf(0.5f * (1.0f + x))) - ORF_PI_LO);
    if (x > 0.5f) return 2.0f * orf_asin_abs(__builtin_sqrtf(0.5f * (1.0f - x)));
    return ORF_PIO2F - (orf_asinf(x) - ORF_PIO2_LO);
}

/* atan (Cephes atanf.c). */
ORF_MATH_FN float orf_atanf(float xx)
{
    if (xx != xx) return xx;
    float x = orf_fabsf(xx), hi = 0.0f, lo = 0.0f;
    if (x > 2.414213562373095f) {
        hi = ORF_PIO2F;
        lo = ORF_PIO2_LO;
        x = -(1.0f / x);
    } else if (x > 0.4142135623730950f) {
        hi = ORF_PIO4F;
        lo = ORF_PIO4_LO;
        x = (x - 1.0f) / (x + 1.0f);
    }
    float z = x * x;
    float w = (((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x +
              x;
    float y = hi + (w + lo);
    return xx < 0.0f ? -y : y;
}

/* atan2 with the C99 quadrant rules for finite arguments and zeros. */
ORF_MATH_FN float orf_atan2f(float y, float x)
{
    if (x != x || y != y) return x + y;
    int yneg = (orf_f2u(y) >> 31) != 0;
    if (x == 0.0f) {
        if (y == 0.0f) { /* atan2(+-0, +0) = +-0, atan2(+-0, -0) = +-pi */
            float r = (orf_f2u(x) >> 31) ? ORF_PIF : 0.0f;
            return yneg ? -r : r;
        }
        return yneg ? -ORF_PIO2F : ORF_PIO2F;
    }
    if (y == 0.0f) {
        float r = x < 0.0f ? ORF_PIF : 0.0f;
        return yneg ? -r : r;
    }
    float z = orf_atanf(y / x);
    if (x < 0.0f) return yneg ? (z - ORF_PI_LO) - ORF_PIF : (z + ORF_PI_LO) + ORF_PIF;
    return z;
}

/* x * 2^n for n in [-300, 300] (two exact power-of-two steps, one rounding). */
ORF_MATH_FN float orf_ldexpf(float x, int n)
{
    int n1 = n / 2, n2 = n - n1;
    if (n1 < -126) n1 = -126;
    if (n1 > 127) n1 = 127;
    if (n2 < -126) n2 = -126;
    if (n2 > 127) n2 = 127;
    return x * orf_u2f((uint32_t)(n1 + 127) << 23) * orf_u2f((uint32_t)(n2 + 127) << 23);
}

/* natural log (Cephes logf.c). */
ORF_MATH_FN float orf_logf(float xx)
{
    if (xx != xx) return xx;
    if (xx < 0.0f) return orf_nanf();
    if (xx == 0.0f) return -orf_u2f(0x7f800000u);
    if (xx == orf_u2f(0x7f800000u)) return xx;
    int e = 0;
    float x = xx;
    if (x < 1.17549435e-38f) { /* subnormal: scale by 2^25 */
        x *= 33554432.0f;
        e = -25;
    }
    uint32_t u = orf_f2u(x);
    e += (int)((u >> 23) & 0xffu) - 126;
    x = orf_u2f((u & 0x007fffffu) | 0x3f000000u); /* mantissa in [0.5, 1) */
    if (x < 0.707106781186547524f) {
        e -= 1;
        x = x + x - 1.0f;
    } else {
        x = x - 1.0f;
    }
    float z = x * x;
    float y = ((((((((7.0376836292e-2f * x - 1.1514610310e-1f) * x + 1.1676998740e-1f) * x - 1.2420140846e-1f) * x +
                   1.4249322787e-1f) * x - 1.6668057665e-1f) * x + 2.0000714765e-1f) * x - 2.4999993993e-1f) * x +
               3.3333331174e-1f) * x * z;
    float fe = (float)e;
    if (e) y += -2.12194440e-4f * fe;
    y += -0.5f * z;
    z = x + y;
    if (e) z += 0.693359375f * fe;
    return z;
}

/* exp (Cephes expf.c). */
ORF_MATH_FN float orf_expf(float xx)
{
    if (xx != xx) return xx;
    if (xx > 88.72283905206835f) return orf_u2f(0x7f800000u);
    if (xx < -103.278929903431851103f) return 0.0f;
    float z = 1.44269504088896341f * xx + 0.5f;
    int n = (int)z;
    if ((float)n > z) n -= 1; /* floor */
    float fn = (float)n;
    float x = xx - fn * 0.693359375f;
    x = x - fn * -2.12194440e-4f;
    z = x * x;
    z = (((((1.9875691500e-4f * x + 1.3981999507e-3f) * x + 8.3334519073e-3f) * x + 4.1665795894e-2f) * x +
          1.6666665459e-1f) * x + 5.0000001201e-1f) * z + x + 1.0f;
    return orf_ldexpf(z, n);
}

/* erf: x P(x^2) below 1, 1 - exp(-x^2) Q((x - 2.46) / 1.46) up to 3.92, 1 above. */
ORF_MATH_FN float orf_erff(float xx)
{
    if (xx != xx) return xx;
    float x = orf_fabsf(xx), r;
    if (x < 1.0f) {
        float z = x * x;
        float p = (((((7.843559433e-05f * z - 8.003825787e-04f) * z + 5.187146831e-03f) * z - 2.685287036e-02f) * z +
                    1.128355116e-01f) * z - 3.761262000e-01f) * z + 1.128379107e+00f;
        r = x * p;
    } else if (x < 3.92f) {
        float u = (x - 2.460000038f) * 6.849315166e-01f;
        float q = 9.884089195e-06f;
        q = q * u - 2.959864469e-05f;
        q = q * u + 5.094023800e-05f;
        q = q * u - 1.284272585e-04f;
        q = q * u + 3.630519495e-04f;
        q = q * u - 9.154251311e-04f;
        q = q * u + 2.210877603e-03f;
        q = q * u - 5.214023404e-03f;
        q = q * u + 1.189966314e-02f;
        q = q * u - 2.616802603e-02f;
        q = q * u + 5.525608733e-02f;
        q = q * u - 1.115168855e-01f;
        q = q * u + 2.138206810e-01f;
        r = 1.0f - orf_expf(-(x * x)) * q;
    } else {
        r = 1.0f;
    }
    return xx < 0.0f ? -r : r;
}

/* x^y = exp(y log x) for x > 0 (not correctly rounded; the same bits on every
 * side), with pow's special cases for x = 0, x = 1, y = 0 and x < 0 (NaN unless
 * y is 0). */
ORF_MATH_FN float orf_powf(float x, float y)
{
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x != x || y != y) return x + y;
    if (x == 0.0f) return y > 0.0f ? 0.0f : orf_u2f(0x7f800000u);
    if (x < 0.0f) return orf_nanf();
    return orf_expf(y * orf_logf(x));
}

#endif /* ORF_MATH_H */
