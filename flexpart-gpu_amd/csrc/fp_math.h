/*
 * fp_math.h - deterministic single-precision transcendentals for the
 * per-particle timestep.
 *
 * The particle step calls exp, log, sin/cos and pow of f32 arguments
 * (Hanna profiles, Box-Muller, Ornstein-Uhlenbeck decay, deposition).
 * Platform libraries (glibc on the host, the ROCm device library on the
 * GPU) round these differently in the last bit, which makes two runs of the
 * same trajectory drift apart through the Hanna regime switches and the PBL
 * reflection.  Every function here is written only with IEEE-754 operations
 * that are correctly rounded on both x86-64 and gfx950 (+, -, *, /, sqrt,
 * fused multiply-add, rint, floor, integer bit manipulation), so the same
 * source gives bit-identical results on the CPU and on the GPU.  The file
 * compiles as C99 (gcc) and as HIP device/host code (hipcc).
 *
 * Both sides must be compiled with floating-point contraction disabled
 * (-ffp-contract=off); every fused multiply-add below is written out
 * explicitly with fmaf/fma.
 *
 * Accuracy (measured against glibc over dense sweeps in the CPU tests):
 *   fpm_expf      <= 1 ulp         (Cody-Waite reduction, degree-7 Taylor)
 *   fpm_logf      <= 1 ulp         (atanh series in s = f/(2+f), 4 terms)
 *   fpm_sincosf   <= 1 ulp on |x| <= 2*pi, the Box-Muller angle range
 *                 (3-part Cody-Waite reduction by pi/2, Taylor kernels)
 *   fpm_powf      correctly rounded except within ~1e-7 ulp of a midpoint
 *                 (exp(y*log(x)) evaluated in binary64, rounded once)
 */
#ifndef FP_MATH_H
#define FP_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define FPM_FN __host__ __device__ static inline
#else
#define FPM_FN static inline
#endif

#define FPM_INF __builtin_inff()
#define FPM_NAN __builtin_nanf("")

FPM_FN uint32_t fpm_f2u(float x) { uint32_t u; __builtin_memcpy(&u, &x, 4); return u; }
FPM_FN float fpm_u2f(uint32_t u) { float x; __builtin_memcpy(&x, &u, 4); return x; }
FPM_FN uint64_t fpm_d2u(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
FPM_FN double fpm_u2d(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }

/* e^x, single precision.  Special cases are selected at the end rather than
 * branched on, so a wavefront never splits inside the function. */
FPM_FN float fpm_expf(float x)
{
    const float xc = __builtin_fminf(__builtin_fmaxf(x, -104.0f), 89.0f);
    const float k = __builtin_rintf(xc * 1.44269502162933349609375f);
    /* ln2 = 0.693145751953125 (exact, 12 trailing zero bits) + 1.4286068e-6 */
    float r = __builtin_fmaf(k, -0.693145751953125f, xc);
    r = __builtin_fmaf(k, -1.428606765330187e-06f, r);
    /* e^r = 1 + r + r^2 * q(r), q from the Taylor series 1/2!, ..., 1/7! */
    float q = 1.98412701e-04f;
    q = __builtin_fmaf(q, r, 1.38888892e-03f);
    q = __builtin_fmaf(q, r, 8.33333377e-03f);
    q = __builtin_fmaf(q, r, 4.16666679e-02f);
    q = __builtin_fmaf(q, r, 1.66666672e-01f);
    q = __builtin_fmaf(q, r, 0.5f);
    const float p = __builtin_fmaf(q, r * r, r);
    /* ldexp rounds y * 2^k once, exactly like fpm_scale2; it overflows to
     * +inf for x in (88.72, 89]. */
    const float res = __builtin_ldexpf(1.0f + p, (int)k);
    return x != x ? x : (x < -104.0f ? 0.0f : res);
}

/* Natural logarithm, single precision (special cases selected, not branched). */
FPM_FN float fpm_logf(float x)
{
    const uint32_t ix0 = fpm_f2u(x);
    const int sub = ix0 < 0x00800000u; /* +0 or subnormal: scale into the normal range */
    const uint32_t ix = sub ? fpm_f2u(x * 8388608.0f) : ix0;
    /* Split x = 2^e * m with m in [sqrt(1/2), sqrt(2)). */
    const uint32_t t = ix - 0x3f3504f3u;
    const int e = (((int32_t)t) >> 23) - (sub ? 23 : 0);
    const float m = fpm_u2f(ix - (t & 0xff800000u));
    const float f = m - 1.0f; /* exact (Sterbenz) */
    const float s = f / (2.0f + f);
    const float z = s * s;
    /* log(1+f) = f - hfsq + s*(hfsq + R), R = 2z/3 + 2z^2/5 + 2z^3/7 + 2z^4/9 */
    float R = 2.22222224e-01f;
    R = __builtin_fmaf(R, z, 2.85714298e-01f);
    R = __builtin_fmaf(R, z, 4.00000006e-01f);
    R = __builtin_fmaf(R, z, 6.66666687e-01f);
    R = R * z;
    const float hfsq = 0.5f * f * f;
    const float dk = (float)e;
    float res = dk * 0.693145751953125f -
                ((hfsq - (s * (hfsq + R) + dk * 1.428606765330187e-06f)) - f);
    res = x == FPM_INF ? x : res;
    res = x == 0.0f ? -FPM_INF : res;
    res = x < 0.0f ? FPM_NAN : res;
    return x != x ? x : res;
}

/* fpm_logf restricted to normal x in (0, 1]: the same arithmetic without the
 * subnormal rescaling and special-value selects, hence the same values on
 * that domain (the Box-Muller radius argument lies in [2^-126, 1)). */
FPM_FN float fpm_logf_unit(float x)
{
    const uint32_t ix = fpm_f2u(x);
    const uint32_t t = ix - 0x3f3504f3u;
    const int e = ((int32_t)t) >> 23;
    const float m = fpm_u2f(ix - (t & 0xff800000u));
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float R = 2.22222224e-01f;
    R = __builtin_fmaf(R, z, 2.85714298e-01f);
    R = __builtin_fmaf(R, z, 4.00000006e-01f);
    R = __builtin_fmaf(R, z, 6.66666687e-01f);
    R = R * z;
    const float hfsq = 0.5f * f * f;
    const float dk = (float)e;
    return dk * 0.693145751953125f -
           ((hfsq - (s * (hfsq + R) + dk * 1.428606765330187e-06f)) - f);
}

/* sin(x) and cos(x), single precision; accurate on |x| <= 2*pi (the first
 * reduction step is exact there), usable with growing error up to |x|~1e4. */
FPM_FN void fpm_sincosf(float x, float *sin_out, float *cos_out)
{
    const float q = __builtin_rintf(x * 0.636619747f);
    /* pi/2 = 1.57079637 - 4.37113883e-08 - 1.71512451e-15 */
    float r = __builtin_fmaf(q, -1.57079637050628662109375f, x);
    r = __builtin_fmaf(q, 4.371138828673793e-08f, r);
    r = __builtin_fmaf(q, 1.7151245100058819e-15f, r);
    const float z = r * r;
    /* sin r = r + r^3 * (-1/3! + z/5! - z^2/7! + z^3/9!) */
    float ps = 2.75573188e-06f;
    ps = __builtin_fmaf(ps, z, -1.98412701e-04f);
    ps = __builtin_fmaf(ps, z, 8.33333377e-03f);
    ps = __builtin_fmaf(ps, z, -1.66666672e-01f);
    const float sr = __builtin_fmaf(r * z, ps, r);
    /* cos r = 1 + z * (-1/2 + z/4! - z^2/6! + z^3/8! - z^4/10!) */
    float pc = -2.75573200e-07f;
    pc = __builtin_fmaf(pc, z, 2.48015876e-05f);
    pc = __builtin_fmaf(pc, z, -1.38888892e-03f);
    pc = __builtin_fmaf(pc, z, 4.16666679e-02f);
    pc = __builtin_fmaf(pc, z, -0.5f);
    const float cr = __builtin_fmaf(z, pc, 1.0f);
    const int quadrant = ((int)q) & 3;
    const float s = (quadrant & 1) ? cr : sr; /* quadrants 1, 3 swap sin and cos */
    const float c = (quadrant & 1) ? sr : cr;
    *sin_out = (quadrant & 2) ? -s : s;       /* quadrants 2, 3 negate sin */
    *cos_out = ((quadrant + 1) & 2) ? -c : c; /* quadrants 1, 2 negate cos */
}

/* ---- binary64 kernels used by fpm_powf ------------------------------- */

/* log(x) for a positive, finite, normal binary64 x; ~1e-16 relative. */
FPM_FN double fpm_log_d(double x)
{
    const uint64_t ix = fpm_d2u(x);
    /* Split x = 2^e * m with m in [sqrt(1/2), sqrt(2)). */
    const uint64_t t = ix - 0x3fe6a09e667f3bcdull;
    const int e = (int)(((int64_t)t) >> 52);
    const double m = fpm_u2d(ix - (t & 0xfff0000000000000ull));
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    /* R = sum_{k=1..8} 2 z^k / (2k+1) */
    double R = 2.0 / 17.0;
    R = __builtin_fma(R, z, 2.0 / 15.0);
    R = __builtin_fma(R, z, 2.0 / 13.0);
    R = __builtin_fma(R, z, 2.0 / 11.0);
    R = __builtin_fma(R, z, 2.0 / 9.0);
    R = __builtin_fma(R, z, 2.0 / 7.0);
    R = __builtin_fma(R, z, 2.0 / 5.0);
    R = __builtin_fma(R, z, 2.0 / 3.0);
    R = R * z;
    const double hfsq = 0.5 * f * f;
    const double dk = (double)e;
    /* ln2 = 0.6931471805598903 (exact, 11 trailing zero bits) + 5.4979e-14 */
    return dk * 0.69314718055989033 -
           ((hfsq - (s * (hfsq + R) + dk * 5.497923018708371e-14)) - f);
}

/* e^t rounded to single precision, for t in [-104, 89]; ~1e-16 relative
 * before the final rounding (overflow and underflow selected at the end). */
FPM_FN float fpm_exp_d_to_f(double t)
{
    const double tc = __builtin_fmin(__builtin_fmax(t, -104.0), 89.0);
    const double k = __builtin_rint(tc * 1.4426950408889634);
    double r = __builtin_fma(k, -0.69314718055989033, tc);
    r = __builtin_fma(k, -5.497923018708371e-14, r);
    /* Taylor series of e^r to degree 11 */
    double p = 1.0 / 39916800.0;
    p = __builtin_fma(p, r, 1.0 / 3628800.0);
    p = __builtin_fma(p, r, 1.0 / 362880.0);
    p = __builtin_fma(p, r, 1.0 / 40320.0);
    p = __builtin_fma(p, r, 1.0 / 5040.0);
    p = __builtin_fma(p, r, 1.0 / 720.0);
    p = __builtin_fma(p, r, 1.0 / 120.0);
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r * r, r);
    const double y = 1.0 + p;
    const int ki = (int)k;
    /* binary64 scaling is exact for ki in [-152, 130] */
    float res = (float)(y * fpm_u2d((uint64_t)(ki + 1023) << 52));
    res = t > 89.0 ? FPM_INF : res;
    return t < -104.0 ? 0.0f : res;
}

/* x^y, single precision, for x >= 0 (negative x gives NaN). */
FPM_FN float fpm_powf(float x, float y)
{
    const float xs = (x > 0.0f && x != FPM_INF) ? x : 1.0f;
    float res = fpm_exp_d_to_f((double)y * fpm_log_d((double)xs));
    /* special cases, lowest priority first */
    res = x == FPM_INF ? (y > 0.0f ? FPM_INF : 0.0f) : res;
    res = x == 0.0f ? (y > 0.0f ? 0.0f : FPM_INF) : res;
    res = x < 0.0f ? FPM_NAN : res;
    res = x == 1.0f ? 1.0f : res;
    res = (x != x || y != y) ? x + y : res;
    return y == 0.0f ? 1.0f : res;
}

FPM_FN float fpm_sqrtf(float x) { return __builtin_sqrtf(x); }

#endif /* FP_MATH_H */
