/*
 * oracle_fmath.h -- transcendental functions of the oracle's R type (TEST INFRASTRUCTURE).
 *
 * FP32 build: the FP32 functions of open_rendering_framework_amd/csrc/orf_math.h
 * (Cephes-method sin/cos/acos/atan2/log/exp and a fitted erf built from IEEE
 * +,-,*,/,sqrt only), so oracle and product round every transcendental the same
 * way and per-sample comparisons measure the restatement, not two libms.  Those
 * functions are pinned separately against double-precision libm (ulp bounds over
 * dense sweeps, tests/test_orf_math.py).
 * FP64 build: double-precision libm, independent of the product.
 */
#ifndef ORACLE_FMATH_H
#define ORACLE_FMATH_H

#include <tgmath.h>

#ifdef ORACLE_FP64
#define F_SIN(x) sin(x)
#define F_COS(x) cos(x)
#define F_TAN(x) tan(x)
#define F_ACOS(x) acos(x)
#define F_ATAN2(y, x) atan2(y, x)
#define F_ATAN(x) atan(x)
#define F_LOG(x) log(x)
#define F_EXP(x) exp(x)
#define F_ERF(x) erf(x)
#define F_POW(x, y) pow(x, y)
#else
#include "../open_rendering_framework_amd/csrc/orf_math.h"
#define F_SIN(x) orf_sinf(x)
#define F_COS(x) orf_cosf(x)
#define F_TAN(x) orf_tanf(x)
#define F_ACOS(x) orf_acosf(x)
#define F_ATAN2(y, x) orf_atan2f(y, x)
#define F_ATAN(x) orf_atanf(x)
#define F_LOG(x) orf_logf(x)
#define F_EXP(x) orf_expf(x)
#define F_ERF(x) orf_erff(x)
#define F_POW(x, y) orf_powf(x, y)
#endif

#endif
