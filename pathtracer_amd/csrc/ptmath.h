/*
 * ptmath.h — portable, deterministic FP32 transcendentals for the spectral tracer.
 *
 * The reference shader (src/shader.comp) calls GLSL.std.450 built-ins whose
 * results are left to the Vulkan driver (sin, cos, exp, log, pow, acos,
 * inversesqrt, normalize, refract, mod, mix ...). This header fixes one
 * definition for each of them, written from textbook methods:
 *
 *   sin/cos : Cody–Waite reduction by pi/2 with a three-part pi/2 and fmaf,
 *             then the Taylor series of sin and cos on [-pi/4, pi/4].
 *   exp     : k = rint(x / ln2), two-part ln2 reduction, Taylor series of e^r
 *             to degree 7, scaling by exact powers of two.
 *   log     : x = m * 2^e with m in [sqrt(1/2), sqrt(2)), s = f / (2 + f),
 *             log(1+f) = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R(s^2))) with the
 *             atanh series R(z) = sum 2 z^k / (2k+1)  (the fdlibm arrangement).
 *   pow     : exp(y * log(x)) evaluated in binary64 (own log/exp series) and
 *             rounded once to binary32.
 *   acos    : binary64 arcsine series with the half-angle identity for |x|>1/2.
 *
 * Only IEEE-754 basic operations (+ - * / sqrt fma floor rint, which are
 * correctly rounded on x86-64 SSE and on gfx950 with correctly-rounded
 * divide/sqrt and denormals kept) are used, so the same source compiled by
 * gcc for the host and by hipcc for the device gives bit-identical results.
 * Both builds must use -ffp-contract=off and no fast-math.
 *
 * The header is C99 and C++/HIP compatible.
 */
#ifndef PTMATH_H
#define PTMATH_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define PTM_FN __host__ __device__ static inline
#else
#include <math.h>
#define PTM_FN static inline
#endif

/* ---- bit helpers ---------------------------------------------------------------- */
PTM_FN uint32_t ptm_f2u(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }
PTM_FN float ptm_u2f(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }
PTM_FN uint64_t ptm_d2u(double f) { uint64_t u; __builtin_memcpy(&u, &f, 8); return u; }
PTM_FN double ptm_u2d(uint64_t u) { double f; __builtin_memcpy(&f, &u, 8); return f; }

#define PTM_INF (ptm_u2f(0x7f800000u))
#define PTM_NAN (ptm_u2f(0x7fc00000u))

PTM_FN int ptm_isnan(float x) { return x != x; }

/* 2^k as a float for -126 <= k <= 127 (exact). */
PTM_FN float ptm_pow2i(int k) { return ptm_u2f((uint32_t)(k + 127) << 23); }
/* 2^k as a double for -1022 <= k <= 1023 (exact). */
PTM_FN double ptm_pow2i_d(int k) { return ptm_u2d((uint64_t)(k + 1023) << 52); }

/* x * 2^k with one rounding at most (the final multiply), for |k| <= 252. */
PTM_FN float ptm_scale2(float x, int k) {
    if (k > 127) return (x * ptm_pow2i(127)) * ptm_pow2i(k - 127);
    if (k < -126) return (x * ptm_pow2i(k + 126)) * ptm_pow2i(-126);
    return x * ptm_pow2i(k);
}

/* ---- basic operations (thin, so every call site reads the same on host and device) */
PTM_FN float ptm_fma(float a, float b, float c) { return fmaf(a, b, c); }
PTM_FN float ptm_sqrt(float x) { return sqrtf(x); }
PTM_FN float ptm_floor(float x) { return floorf(x); }
PTM_FN float ptm_ceil(float x) { return ceilf(x); }
PTM_FN float ptm_abs(float x) { return fabsf(x); }
PTM_FN float ptm_min(float a, float b) { return fminf(a, b); }
PTM_FN float ptm_max(float a, float b) { return fmaxf(a, b); }
PTM_FN float ptm_clamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
/* GLSL sign(): +1, -1, or 0 (NaN maps to 0). */
PTM_FN float ptm_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
/* GLSL step(edge, x). */
PTM_FN float ptm_step(float edge, float x) { return x < edge ? 0.0f : 1.0f; }
/* GLSL mix(x, y, a) = x * (1 - a) + y * a (the specification's formula). */
PTM_FN float ptm_mix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
/* GLSL mod(x, y) = x - y * floor(x / y). */
PTM_FN float ptm_mod(float x, float y) { return x - y * floorf(x / y); }
/* GLSL inversesqrt(x), defined here as a correctly rounded 1 / sqrt(x). */
PTM_FN float ptm_rsqrt(float x) { return 1.0f / sqrtf(x); }

/* ---- sin / cos ------------------------------------------------------------------ */
#define PTM_PIO2_HI 1.5707963705062866f
#define PTM_PIO2_MID (-4.371138828673793e-08f)
#define PTM_PIO2_LO (-1.7151245100058819e-15f)
#define PTM_2OPI 0.6366197466850281f

/* Taylor coefficients 1/n! with alternating signs. */
PTM_FN float ptm__sin_kernel(float r) {
    float z = r * r;
    float p = fmaf(z, 2.7557319e-6f /* 1/9! */, -1.9841270e-4f /* -1/7! */);
    p = fmaf(p, z, 8.3333333e-3f /* 1/5! */);
    p = fmaf(p, z, -1.6666667e-1f /* -1/3! */);
    return fmaf(r * z, p, r);
}
PTM_FN float ptm__cos_kernel(float r) {
    float z = r * r;
    float p = fmaf(z, -2.7557319e-7f /* -1/10! */, 2.4801587e-5f /* 1/8! */);
    p = fmaf(p, z, -1.3888889e-3f /* -1/6! */);
    p = fmaf(p, z, 4.1666667e-2f /* 1/4! */);
    p = fmaf(p, z, -0.5f);
    return fmaf(z, p, 1.0f);
}
/* Reduces x to r in [-pi/4, pi/4] and returns the quadrant in *q. */
PTM_FN float ptm__reduce_pio2(float x, int* q) {
    float k = rintf(x * PTM_2OPI);
    float r = fmaf(-k, PTM_PIO2_HI, x);
    r = fmaf(-k, PTM_PIO2_MID, r);
    r = fmaf(-k, PTM_PIO2_LO, r);
    /* quadrant = k mod 4, exact for every integral float k */
    *q = (int)(k - 4.0f * floorf(k * 0.25f));
    return r;
}
PTM_FN float ptm_sin(float x) {
    if (!(fabsf(x) < PTM_INF)) return PTM_NAN;
    int q;
    float r = ptm__reduce_pio2(x, &q);
    switch (q & 3) {
        case 0: return ptm__sin_kernel(r);
        case 1: return ptm__cos_kernel(r);
        case 2: return -ptm__sin_kernel(r);
        default: return -ptm__cos_kernel(r);
    }
}
PTM_FN float ptm_cos(float x) {
    if (!(fabsf(x) < PTM_INF)) return PTM_NAN;
    int q;
    float r = ptm__reduce_pio2(x, &q);
    switch (q & 3) {
        case 0: return ptm__cos_kernel(r);
        case 1: return -ptm__sin_kernel(r);
        case 2: return -ptm__cos_kernel(r);
        default: return ptm__sin_kernel(r);
    }
}

/* ---- exp ------------------------------------------------------------------------ */
#define PTM_LOG2E 1.4426950216293335f
#define PTM_LN2_HI 0.6931471824645996f
#define PTM_LN2_LO (-1.9046542121259336e-09f)

PTM_FN float ptm_exp(float x) {
    if (x != x) return x;
    if (x > 88.72284f) return PTM_INF;
    if (x < -103.97209f) return 0.0f;
    float k = rintf(x * PTM_LOG2E);
    float r = fmaf(-k, PTM_LN2_HI, x);
    r = fmaf(-k, PTM_LN2_LO, r);
    /* e^r = sum r^n / n!, n <= 7, |r| <= ln2/2 */
    float p = fmaf(r, 1.9841270e-4f /* 1/7! */, 1.3888889e-3f /* 1/6! */);
    p = fmaf(p, r, 8.3333333e-3f);
    p = fmaf(p, r, 4.1666667e-2f);
    p = fmaf(p, r, 1.6666667e-1f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ptm_scale2(p, (int)k);
}

/* ---- log ------------------------------------------------------------------------ */
PTM_FN float ptm_log(float x) {
    if (x != x) return x;
    if (x < 0.0f) return PTM_NAN;
    if (x == 0.0f) return -PTM_INF;
    if (x == PTM_INF) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x *= 33554432.0f; e = -25; } /* denormal: scale by 2^25 */
    uint32_t u = ptm_f2u(x);
    e += (int)((u >> 23) & 0xffu) - 127;
    float m = ptm_u2f((u & 0x007fffffu) | 0x3f800000u); /* [1, 2) */
    if (m > 1.41421354f) { m *= 0.5f; e += 1; }         /* [sqrt(1/2), sqrt(2)] */
    float f = m - 1.0f;                                  /* exact */
    float s = f / (2.0f + f);
    float z = s * s;
    /* R(z) = 2z/3 + 2z^2/5 + ... + 2z^5/11 */
    float R = fmaf(z, 0.18181818f /* 2/11 */, 0.22222222f /* 2/9 */);
    R = fmaf(R, z, 0.28571429f /* 2/7 */);
    R = fmaf(R, z, 0.4f /* 2/5 */);
    R = fmaf(R, z, 0.66666667f /* 2/3 */);
    R = R * z;
    float hfsq = 0.5f * f * f;
    float ef = (float)e;
    return ef * PTM_LN2_HI + (f - (hfsq - (s * (hfsq + R) + ef * PTM_LN2_LO)));
}

/* ---- binary64 helpers used by pow and acos ---------------------------------------- */
#define PTM_LN2_D 0.6931471805599453
#define PTM_LN2_D_LO 2.3190468138462996e-17
#define PTM_LOG2E_D 1.4426950408889634

/* log(x) for finite x > 0 (normal binary64). */
PTM_FN double ptm__log_d(double x) {
    uint64_t u = ptm_d2u(x);
    int e = (int)((u >> 52) & 0x7ffu) - 1023;
    double m = ptm_u2d((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) { m *= 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = fma(z, 2.0 / 25.0, 2.0 / 23.0);
    R = fma(R, z, 2.0 / 21.0);
    R = fma(R, z, 2.0 / 19.0);
    R = fma(R, z, 2.0 / 17.0);
    R = fma(R, z, 2.0 / 15.0);
    R = fma(R, z, 2.0 / 13.0);
    R = fma(R, z, 2.0 / 11.0);
    R = fma(R, z, 2.0 / 9.0);
    R = fma(R, z, 2.0 / 7.0);
    R = fma(R, z, 2.0 / 5.0);
    R = fma(R, z, 2.0 / 3.0);
    R = R * z;
    double hfsq = 0.5 * f * f;
    double ed = (double)e;
    return ed * PTM_LN2_D + (f - (hfsq - (s * (hfsq + R) + ed * PTM_LN2_D_LO)));
}

/* e^x for -745 < x < 709 (result a normal binary64 whenever |x| < 708). */
PTM_FN double ptm__exp_d(double x) {
    double k = rint(x * PTM_LOG2E_D);
    double r = fma(-k, PTM_LN2_D, x);
    r = fma(-k, PTM_LN2_D_LO, r);
    /* Taylor series to degree 13 on |r| <= ln2/2 */
    double p = fma(r, 1.0 / 6227020800.0, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return p * ptm_pow2i_d((int)k);
}

/* ---- pow ------------------------------------------------------------------------ */
/* GLSL pow(x, y): undefined for x < 0 (NaN here), x == 0 with y <= 0 (+inf here). */
PTM_FN float ptm_pow(float x, float y) {
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x != x || y != y) return PTM_NAN;
    if (x < 0.0f) return PTM_NAN;
    if (x == 0.0f) return y > 0.0f ? 0.0f : PTM_INF;
    if (x == PTM_INF) return y > 0.0f ? PTM_INF : 0.0f;
    if (fabsf(y) == PTM_INF) {
        if (x < 1.0f) return y > 0.0f ? 0.0f : PTM_INF;
        return y > 0.0f ? PTM_INF : 0.0f;
    }
    double t = (double)y * ptm__log_d((double)x);
    if (t > 89.0) return PTM_INF;
    if (t < -104.0) return 0.0f;
    return (float)ptm__exp_d(t);
}

/* ---- acos ----------------------------------------------------------------------- */
/* asin(z) for |z| <= 1/2: sum_n (2n)! / (4^n (n!)^2 (2n+1)) z^(2n+1), 30 terms. */
PTM_FN double ptm__asin_series_d(double z) {
    double z2 = z * z, term = z, sum = z, c = 1.0;
    for (int n = 1; n < 30; ++n) {
        c = c * (double)(2 * n - 1) / (double)(2 * n); /* (2n)! / (4^n (n!)^2) */
        term = term * z2;
        sum = fma(c / (double)(2 * n + 1), term, sum);
    }
    return sum;
}
#define PTM_PI_D 3.141592653589793
#define PTM_PIO2_D 1.5707963267948966
PTM_FN float ptm_acos(float xf) {
    if (xf != xf) return xf;
    double x = (double)xf, a = fabs(x);
    if (a > 1.0) return PTM_NAN;
    double r;
    if (a <= 0.5) {
        r = PTM_PIO2_D - ptm__asin_series_d(x);
    } else {
        double h = 2.0 * ptm__asin_series_d(sqrt((1.0 - a) * 0.5));
        r = x > 0.0 ? h : PTM_PI_D - h;
    }
    return (float)r;
}

#endif /* PTMATH_H */
