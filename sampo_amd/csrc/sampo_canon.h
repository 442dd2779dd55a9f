/*
 * sampo_canon.h -- the canonical numeric contract shared by the HIP kernels
 * and the CPU oracle.
 *
 * The reference kernels leave several results implementation-defined:
 * OpenCL's exp/log/pow/sin/sqrt may be approximate, `/` may be 2.5 ulp, the
 * AMD OpenCL compiler contracts a*b+c into FMA, and float->uint conversion of
 * negative, huge or NaN values differs between x86 (wraps) and gfx9
 * (saturates).  This header fixes one answer for each so that host and device
 * produce identical bits:
 *
 *   - binary32 + - * / are IEEE correctly rounded; both sides are compiled
 *     with -ffp-contract=off, so no expression is fused.
 *   - exp/log/sin are evaluated in binary64 from published methods (Cody-Waite
 *     argument reduction with fdlibm's split constants, Taylor/atanh series
 *     evaluated by Horner) using only + - * / and integer bit manipulation,
 *     then rounded once to binary32.  Identical operation sequences give
 *     identical bits on any IEEE machine.
 *   - sqrtf is correctly rounded: a binary64 estimate is corrected against
 *     the exact squares of the neighbouring rounding midpoints.
 *   - float->u32 saturates: NaN -> 0, x <= -1 -> 0, x >= 2^32 -> 2^32-1.
 *   - the hybrid Tausworthe/LCG generator and the glibc TYPE_3 additive
 *     feedback generator (rand/random_r) are restated bit-exactly.
 *
 * Written in the common subset of C99 and HIP C++ so that gcc (oracle, host
 * library) and hipcc (device code) compile the same text.
 */
#ifndef SAMPO_CANON_H
#define SAMPO_CANON_H

#include <stdint.h>

#if defined(__HIPCC__)
#define SC_FN __host__ __device__ static inline
#else
#define SC_FN static inline
#endif

/* ------------------------------------------------------------------ bits */

SC_FN uint64_t sc_bits_d(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
SC_FN double sc_double(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }
SC_FN uint32_t sc_bits_f(float x) { uint32_t u; __builtin_memcpy(&u, &x, 4); return u; }
SC_FN float sc_float(uint32_t u) { float x; __builtin_memcpy(&x, &u, 4); return x; }

#define SC_INF_D sc_double(0x7ff0000000000000ull)
#define SC_NAN_D sc_double(0x7ff8000000000000ull)
#define SC_INF_F sc_float(0x7f800000u)
#define SC_NAN_F sc_float(0x7fc00000u)

/* ln2 split (fdlibm): the high part has 21 trailing zero bits, so k*LN2_HI
 * is exact for |k| < 2^21. */
#define SC_LN2_HI sc_double(0x3fe62e42fee00000ull)
#define SC_LN2_LO sc_double(0x3dea39ef35793c76ull)
#define SC_INV_LN2 1.4426950408889634
#define SC_SQRT2 1.4142135623730951

/* multiply p by 2^k for any k that keeps the product a binary64 number or
 * underflows it to zero; two steps keep the scale factor itself normal. */
SC_FN double sc_scale2(double p, int k)
{
	if (k > 1023) {
		p = p * sc_double((uint64_t)(1023 + 1023) << 52);
		k -= 1023;
		if (k > 1023) k = 1024; /* overflow to inf below */
	}
	if (k < -1022) {
		p = p * sc_double((uint64_t)(1) << 52); /* 2^-1022 */
		k += 1022;
		if (k < -1022) return 0.0 * p;
	}
	if (k == 1024) return p * SC_INF_D;
	return p * sc_double((uint64_t)(k + 1023) << 52);
}

/* ------------------------------------------------------------- exp (f64)
 * Cody-Waite reduction x = k*ln2 + r, |r| <= ln2/2, then the Taylor series of
 * e^r to degree 13 (truncation < 5e-18) by Horner, then scaling by 2^k. */
SC_FN double sc_exp_d(double x)
{
	if (x != x) return x;
	if (x > 709.782712893384) return SC_INF_D;
	if (x < -745.2) return 0.0;
	double t = x * SC_INV_LN2;
	int k = (int)(t < 0.0 ? t - 0.5 : t + 0.5);
	double kd = (double)k;
	double r = (x - kd * SC_LN2_HI) - kd * SC_LN2_LO;
	double p = 1.0 / 6227020800.0;        /* 1/13! */
	p = p * r + 1.0 / 479001600.0;        /* 1/12! */
	p = p * r + 1.0 / 39916800.0;
	p = p * r + 1.0 / 3628800.0;
	p = p * r + 1.0 / 362880.0;
	p = p * r + 1.0 / 40320.0;
	p = p * r + 1.0 / 5040.0;
	p = p * r + 1.0 / 720.0;
	p = p * r + 1.0 / 120.0;
	p = p * r + 1.0 / 24.0;
	p = p * r + 1.0 / 6.0;
	p = p * r + 0.5;
	p = p * r + 1.0;
	p = p * r + 1.0;
	return sc_scale2(p, k);
}

/* ------------------------------------------------------------- log (f64)
 * x = 2^e * m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
 * s = (m-1)/(m+1), |s| <= 0.1716, series to s^23 (truncation < 1e-18). */
SC_FN double sc_log_d(double x)
{
	if (x != x) return x;
	if (x < 0.0) return SC_NAN_D;
	if (x == 0.0) return -SC_INF_D;
	if (x == SC_INF_D) return x;
	uint64_t u = sc_bits_d(x);
	int e = (int)((u >> 52) & 0x7ff);
	if (e == 0) { /* subnormal: scale by 2^54 */
		x = x * sc_double((uint64_t)(1023 + 54) << 52);
		u = sc_bits_d(x);
		e = (int)((u >> 52) & 0x7ff) - 54;
	}
	e -= 1023;
	double m = sc_double((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
	if (m > SC_SQRT2) { m = m * 0.5; e += 1; }
	double f = m - 1.0;
	double s = f / (2.0 + f);
	double z = s * s;
	double q = 1.0 / 23.0;
	q = q * z + 1.0 / 21.0;
	q = q * z + 1.0 / 19.0;
	q = q * z + 1.0 / 17.0;
	q = q * z + 1.0 / 15.0;
	q = q * z + 1.0 / 13.0;
	q = q * z + 1.0 / 11.0;
	q = q * z + 1.0 / 9.0;
	q = q * z + 1.0 / 7.0;
	q = q * z + 1.0 / 5.0;
	q = q * z + 1.0 / 3.0;
	q = q * z + 1.0;
	double lm = (2.0 * s) * q;
	double ed = (double)e;
	return ed * SC_LN2_HI + (ed * SC_LN2_LO + lm);
}

/* ------------------------------------------------------------- sin (f64)
 * x = k*pi/2 + r with pi/2 split in three 33-bit pieces (fdlibm), exact
 * products for |k| < 2^20; quadrant selects +-sin r / +-cos r, each a Taylor
 * polynomial on |r| <= pi/4 (truncation < 1e-21). */
SC_FN double sc_sin_d(double x)
{
	if (x != x || x == SC_INF_D || x == -SC_INF_D) return SC_NAN_D;
	const double pio2_1 = sc_double(0x3ff921fb54400000ull);
	const double pio2_2 = sc_double(0x3dd0b4611a600000ull);
	const double pio2_2t = sc_double(0x3ba3198a2e037073ull);
	double t = x * 0.63661977236758134; /* 2/pi */
	int k = (int)(t < 0.0 ? t - 0.5 : t + 0.5);
	double kd = (double)k;
	double r = ((x - kd * pio2_1) - kd * pio2_2) - kd * pio2_2t;
	double z = r * r;
	double sp = -1.0 / 121645100408832000.0; /* -1/19! */
	sp = sp * z + 1.0 / 355687428096000.0;   /* 1/17! */
	sp = sp * z - 1.0 / 1307674368000.0;     /* 1/15! */
	sp = sp * z + 1.0 / 6227020800.0;
	sp = sp * z - 1.0 / 39916800.0;
	sp = sp * z + 1.0 / 362880.0;
	sp = sp * z - 1.0 / 5040.0;
	sp = sp * z + 1.0 / 120.0;
	sp = sp * z - 1.0 / 6.0;
	sp = sp * z + 1.0;
	double sn = r * sp;
	double cp = 1.0 / 2432902008176640000.0; /* 1/20! */
	cp = cp * z - 1.0 / 6402373705728000.0;  /* 1/18! */
	cp = cp * z + 1.0 / 20922789888000.0;    /* 1/16! */
	cp = cp * z - 1.0 / 87178291200.0;       /* 1/14! */
	cp = cp * z + 1.0 / 479001600.0;
	cp = cp * z - 1.0 / 3628800.0;
	cp = cp * z + 1.0 / 40320.0;
	cp = cp * z - 1.0 / 720.0;
	cp = cp * z + 1.0 / 24.0;
	cp = cp * z - 0.5;
	cp = cp * z + 1.0;
	switch (k & 3) {
	case 0: return sn;
	case 1: return cp;
	case 2: return -sn;
	default: return -cp;
	}
}

/* ----------------------------------------------------- binary32 functions */

SC_FN float sc_expf(float x) { return (float)sc_exp_d((double)x); }
SC_FN float sc_logf(float x) { return (float)sc_log_d((double)x); }
SC_FN float sc_sinf(float x) { return (float)sc_sin_d((double)x); }

/* correctly rounded square root: no binary32 value is the exact square root
 * of a midpoint square, so comparing x with the exact (<= 50-bit) squares of
 * the two neighbouring midpoints decides the rounding without ties. */
SC_FN float sc_sqrtf(float x)
{
	if (x != x || x < 0.0f) return SC_NAN_F;
	if (x == 0.0f || x == SC_INF_F) return x;
	double xd = (double)x;
	float r = (float)__builtin_sqrt(xd);
	uint32_t rb = sc_bits_f(r);
	float up = sc_float(rb + 1u);
	float dn = sc_float(rb - 1u);
	double mhi = ((double)r + (double)up) * 0.5;
	double mlo = ((double)r + (double)dn) * 0.5;
	if (xd > mhi * mhi) r = up;
	else if (xd < mlo * mlo) r = dn;
	return r;
}

/* pow(x, y) = exp(y log x) in binary64, rounded once. */
SC_FN float sc_powf(float x, float y)
{
	if (y == 0.0f) return 1.0f;
	if (x == 1.0f) return 1.0f;
	if (x != x || y != y) return SC_NAN_F;
	if (x > 0.0f) return (float)sc_exp_d((double)y * sc_log_d((double)x));
	if (x == 0.0f) return y > 0.0f ? 0.0f : SC_INF_F;
	/* x < 0: defined for integral y only */
	float yi = __builtin_truncf(y);
	if (yi != y) return SC_NAN_F;
	float m = (float)sc_exp_d((double)y * sc_log_d(-(double)x));
	double half = (double)y * 0.5;
	int odd = (__builtin_trunc(half) != half);
	return odd ? -m : m;
}

SC_FN float sc_ceilf(float x) { return __builtin_ceilf(x); }

/* float -> uint32 with saturation (gfx9 hardware semantics). */
SC_FN uint32_t sc_f2u(float x)
{
	if (!(x > -1.0f)) return 0u; /* NaN and x <= -1 */
	if (x >= 4294967296.0f) return 0xffffffffu;
	if (x < 1.0f) return 0u;
	return (uint32_t)x;
}

/* OpenCL C min/max for floats: min returns y if y < x, else x;
 * max returns y if x < y, else x. */
SC_FN float sc_minf(float x, float y) { return (y < x) ? y : x; }
SC_FN float sc_maxf(float x, float y) { return (x < y) ? y : x; }

/* -------------------------------------------------- device random numbers
 * Hybrid Tausworthe / LCG generator (GPU Gems 3 ch. 37).  Stateless: one
 * draw from seeds (z1..z4) for global work-item id gid.  The scale constant
 * 2.3283064365387e-10f is exactly 2^-32 (0x2f800000). */
SC_FN uint32_t sc_taus_step(uint32_t z, int s1, int s2, int s3, uint32_t m)
{
	uint32_t b = ((z << s1) ^ z) >> s2;
	return ((z & m) << s3) ^ b;
}

SC_FN uint32_t sc_hybrid_taus_bits(uint32_t z1, uint32_t z2, uint32_t z3, uint32_t z4, uint32_t gid)
{
	uint32_t a = z1 + 128u;
	uint32_t b = z2 + 128u + gid;
	uint32_t c = z3 + 128u;
	uint32_t d = z4 * gid;
	return sc_taus_step(a, 13, 19, 12, 0xfffffffeu) ^
	       sc_taus_step(b, 2, 25, 4, 0xfffffff8u) ^
	       sc_taus_step(c, 3, 11, 17, 0xfffffff0u) ^
	       (1664525u * d + 1013904223u);
}

SC_FN float sc_hybrid_taus(uint32_t z1, uint32_t z2, uint32_t z3, uint32_t z4, uint32_t gid)
{
	return sc_float(0x2f800000u) * (float)sc_hybrid_taus_bits(z1, z2, z3, z4, gid);
}

/* Box-Muller transform on two uniforms, binary32, evaluated in the order the
 * reference writes it: r = sqrt(-2 * log(u0)); theta = (2 * pi_f) * u1;
 * (r * sin(theta)) * sd + mean. */
#define SC_PI_F 3.14159274101257324e+00f
SC_FN float sc_box_muller(float u0, float u1, float mean, float sd)
{
	float r = sc_sqrtf(-2.0f * sc_logf(u0));
	float theta = (2.0f * SC_PI_F) * u1;
	return (r * sc_sinf(theta)) * sd + mean;
}

/* ------------------------------------------------ host seed stream (glibc)
 * glibc's rand() is random_r over a TYPE_3 additive lagged-Fibonacci state
 * of 31 words (x[i] = x[i-3] + x[i-31] mod 2^32, output x >> 1), seeded by
 * the Park-Miller minimal standard LCG (16807, 2^31-1) and advanced 310 times
 * before the first output. */
typedef struct {
	int32_t st[31];
	uint32_t f, r;
} sc_glibc_rand;

SC_FN uint32_t sc_glibc_next(sc_glibc_rand *g)
{
	uint32_t v = (uint32_t)g->st[g->f] + (uint32_t)g->st[g->r];
	g->st[g->f] = (int32_t)v;
	g->f = (g->f + 1u == 31u) ? 0u : g->f + 1u;
	g->r = (g->r + 1u == 31u) ? 0u : g->r + 1u;
	return v >> 1;
}

SC_FN void sc_glibc_seed(sc_glibc_rand *g, uint32_t seed)
{
	if (seed == 0u) seed = 1u;
	int64_t w = (int64_t)(int32_t)seed;
	g->st[0] = (int32_t)seed;
	for (int i = 1; i < 31; ++i) {
		/* Schrage's form of 16807 * w mod (2^31 - 1) */
		int64_t hi = w / 127773;
		int64_t lo = w % 127773;
		w = 16807 * lo - 2836 * hi;
		if (w < 0) w += 2147483647;
		g->st[i] = (int32_t)w;
	}
	g->f = 3u;
	g->r = 0u;
	for (int i = 0; i < 310; ++i) (void)sc_glibc_next(g);
}

#endif /* SAMPO_CANON_H */
