// fp_physics.hpp - device functions of the per-particle timestep shared by
// the forward step kernel (fp_step.hip) and the adjoint (fp_adj.hip): Philox,
// met sampling, Hanna profiles, the Langevin updates, reflection and the
// gridding helpers.  Each source file that includes it gets its own copy
// (anonymous namespace); the forward kernel and the adjoint's recomputation
// of a step therefore evaluate the same expressions.
#pragma once

#include "fp_internal.hpp"
#include "fp_math.h"

namespace fp {
namespace {

constexpr uint32_t FLAG_ACTIVE = 1u;
constexpr float USTAR_MIN = 1.0e-4f;
constexpr float TLU_MIN = 10.0f;
constexpr float TLV_MIN = 10.0f;
constexpr float TLW_MIN = 30.0f;
constexpr float SIGMA_FLOOR = 1.0e-2f;
constexpr float H_MIN = 1.0f;
constexpr float F32_MAX = 3.40282347e38f;

__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
// clampf of a value produced by arithmetic (never a signaling NaN), lo <= hi,
// as one median-of-three: for a quiet NaN x, v_med3_f32 returns the minimum
// of its operands that are numbers, lo, which is what fminf(fmaxf(x, lo), hi)
// gives; otherwise both are the middle value.  One instruction instead of a
// max and a min.
__device__ __forceinline__ float clampq(float x, float lo, float hi) { return __builtin_amdgcn_fmed3f(x, lo, hi); }
// Floating-point class tests (v_cmp_class_f32): bit 7 positive subnormal, bit 8
// positive normal; 0x1f8 = every finite class (normals, subnormals, zeros).
__device__ __forceinline__ bool finite_f(float v) { return __builtin_amdgcn_classf(v, 0x1f8); }
__device__ __forceinline__ float pos_or_zero(float v) { return __builtin_amdgcn_classf(v, 0x180) ? v : 0.0f; }
// a*(1-t) + b*t with the product-sum fused, as the reference CPU twin does
// (a.mul_add(1.0 - t, b * t), physics/interpolation.rs:76-79); u = 1 - t.
__device__ __forceinline__ float mix(float a, float b, float t, float u) { return __builtin_fmaf(a, u, b * t); }

// Correctly rounded square root of x in [2^-96, FLT_MAX]: the hardware
// square root s0 and reciprocal square root y of x, then one correction
// s0 + (x - s0^2) * (y / 2) with the residual x - s0^2 exact in a fused
// multiply-add.  Every float in that range gives the IEEE result; this was
// checked exhaustively on gfx950 by an exact binary64 test of the rounding
// (fp_math_sweep).  Below 2^-96 the residual underflows and the result can be
// off by one unit in the last place.
__device__ __forceinline__ float sqrt_in_range(float x, float *half_rsqrt = nullptr)
{
    const float s0 = __builtin_amdgcn_sqrtf(x);
    const float y = __builtin_amdgcn_rsqf(x);
    const float d = __builtin_fmaf(-s0, s0, x);
    const float hy = 0.5f * y;
    if (half_rsqrt) *half_rsqrt = hy;  // about 1 / (2 sqrt(x))
    return __builtin_fmaf(d, hy, s0);
}

// Correctly rounded square root of any x: a wavefront with a lane outside
// [2^-96, FLT_MAX] (zero, tiny, negative, infinite or NaN) takes the
// compiler's IEEE sequence.
__device__ __forceinline__ float sqrt_cr(float x)
{
    if (__builtin_expect(__ballot(!(x >= 0x1p-96f && x <= F32_MAX)) != 0ull, 0)) return __builtin_sqrtf(x);
    return sqrt_in_range(x);
}

// sqrt_cr of an x known not to be +infinity: only the lower bound is tested.
// The step's Box-Muller radii (-2 log u, u in [2^-126, 1)) and
// Ornstein-Uhlenbeck noise scales (arguments in [0, 1]) are such arguments.
__device__ __forceinline__ float sqrt_cr_finite(float x)
{
    if (__builtin_expect(__ballot(!(x >= 0x1p-96f)) != 0ull, 0)) return __builtin_sqrtf(x);
    return sqrt_in_range(x);
}

// ---------------------------------------------------------------- Philox

struct U4 { uint32_t x, y, z, w; };

// a ^ b ^ c in one VALU: gfx950's three-input bit operation with truth table
// 0x96 (the parity of its three inputs); CDNA has no v_xor3_b32.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ U4 philox10(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 multiply per lane and word (v_mad_u64_u32)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = U4{xor3((uint32_t)(p1 >> 32), c.y, k0), (uint32_t)p1, xor3((uint32_t)(p0 >> 32), c.w, k1),
               (uint32_t)p0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// base + offset, 128-bit wrapping (physics/rng.rs:69-87)
__device__ __forceinline__ U4 ctr_add(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t off)
{
    const uint64_t lo = (uint64_t)c0 + (uint32_t)off;
    const uint64_t mid = (uint64_t)c1 + (uint32_t)(off >> 32) + (lo >> 32);
    const uint64_t hi = (uint64_t)c2 + (mid >> 32);
    return U4{(uint32_t)lo, (uint32_t)mid, (uint32_t)hi, c3 + (uint32_t)(hi >> 32)};
}

__device__ __forceinline__ float unit24(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }

// box_muller_normals (langevin_fused.wgsl:186-192)
__device__ __forceinline__ void gauss_pair(uint32_t a, uint32_t b, float &g0, float &g1)
{
    // unit24 lies in [0, 1 - 2^-24] and 0.99999994f is 1 - 2^-24, so the
    // reference's upper clamps (and the lower one of au) leave it unchanged
    const float ru = fmaxf(unit24(a), 1.1754944e-38f);
    const float au = unit24(b);
    // ru in [2^-126, 1 - 2^-24]: -2 log ru lies in [1.19e-7, 174.7] (every one
    // of the 2^24 values checked on the CPU), inside sqrt_in_range's range
    const float rad = sqrt_in_range(-2.0f * fpm_logf_unit(ru));
    float sn, cs;
    fpm_sincosf(6.2831855f * au, &sn, &cs);
    g0 = rad * cs;
    g1 = rad * sn;
}

// ---------------------------------------------------------------- met sampling

struct Axis { uint32_t lo, hi; float t; };

// coord from arithmetic (clampq); axis_of<true> for a coordinate that may be
// a value loaded from memory as it is (clampf)
// nm = (float)(n - 1), precomputed for the launch
template <bool RAW = false>
__device__ __forceinline__ Axis axis_of(float coord, uint32_t n, float nm)
{
    const float c = RAW ? clampf(coord, 0.0f, nm) : clampq(coord, 0.0f, nm);
    const float fl = floorf(c);
    const uint32_t l = (uint32_t)fl;
    return Axis{l, min(l + 1u, n - 1u), c - fl};
}
template <bool RAW = false>
__device__ __forceinline__ Axis axis_of(float coord, uint32_t n)
{
    return axis_of<RAW>(coord, n, (float)(n - 1));
}

// z [m] -> fractional level: the first k >= 1 with z <= h[k] (the linear scan
// of advection_dual_wind.wgsl:94-117) over the non-decreasing level table in
// LDS.  The search starts at the first level that can qualify for z's height
// band (the table's band index) and steps up from there: every level below
// the band's start is lower than z, so the result is that of the full scan.
__device__ __forceinline__ float level_of_scan(float z, const float *lh, uint32_t nz)
{
    const float h0 = lh[0];
    if (z <= h0) return h0 > 0.0f ? z / h0 : 0.0f;
    const float inv_w = lh[nz + 2];
    const uint16_t *first = reinterpret_cast<const uint16_t *>(lh + nz + 3);
    const uint32_t b = (uint32_t)clampq(z * inv_w, 0.0f, (float)(kLevelBuckets - 1u));
    uint32_t lo = first[b];
    while (lo < nz && !(z <= lh[lo])) ++lo;
    if (lo >= nz) return (float)(nz - 1);
    const float hp = lh[lo - 1];
    const float den = lh[lo] - hp;
    return den > 0.0f ? (float)(lo - 1) + (z - hp) / den : (float)(lo - 1);
}

// The same lookup for a table with h0 > 0, at least two levels and every
// band's first candidate f below the top level (kLevelHeightsBanded): the
// heights are non-decreasing, so the levels below z among f and f + 1 are a
// prefix and two comparisons give the scan's answer whenever it lies within
// two levels of f (the words past the top level are +inf).  A lane whose
// answer lies further up (a band that holds more levels, or z NaN) continues
// the scan from there.  z <= h0 gives f = 1 and the answer 1, so one
// division serves both cases: (z - 0) / (h0 - 0) = z / h0, and -0 + q = q.
// With non-decreasing heights the scan's den = h[lo] - h[lo - 1] is positive
// whenever it stops at lo < nz, so its den > 0 test cannot fail here.
__device__ __forceinline__ float level_of_banded(float z, const float *lh, uint32_t nz)
{
    const float h0 = lh[0];
    const bool below = z <= h0;
    const float inv_w = lh[nz + 2];
    const uint16_t *first = reinterpret_cast<const uint16_t *>(lh + nz + 3);
    const uint32_t b = (uint32_t)clampq(z * inv_w, 0.0f, (float)(kLevelBuckets - 1u));
    const uint32_t f = first[b];
    const float c0 = lh[f], c1 = lh[f + 1u];
    uint32_t lo = f + (!(z <= c0) ? 1u : 0u) + (!(z <= c1) ? 1u : 0u);
    float hp = lh[lo - 1u], hh = lh[lo];
    if (!(z <= hh)) {
        while (lo < nz && !(z <= lh[lo])) ++lo;
        hp = lh[lo - 1u];
        hh = lh[lo];
    }
    const float lp = below ? 0.0f : hp;
    const float lt = below ? hp : hh;
    const float q = (z - lp) / (lt - lp);
    const float base = below ? -0.0f : (float)(lo - 1u);
    return lo < nz ? base + q : (float)(nz - 1u);
}

// heights: kLevelIndex (z is a level index already), kLevelHeights or
// kLevelHeightsBanded (the table's lookup, identical results)
__device__ __forceinline__ float level_of(float z, const float *lh, uint32_t nz, uint32_t heights)
{
    if (heights == kLevelIndex) return z;
    return heights == kLevelHeightsBanded ? level_of_banded(z, lh, nz) : level_of_scan(z, lh, nz);
}

struct Wind { float u, v, w; };

// sample_dual_u/v/w (advection_dual_wind.wgsl:142-236): trilinear per time
// level, then the temporal blend.  The 8 corners come from packed 32-byte
// records {u0,v0,w0,u1,v1,w1,-,-}: two 16-byte loads per corner.
__device__ __forceinline__ Wind sample_wind(const StepArgs &a, const float *lh, float alpha, float x,
                                            float y, float z)
{
    const Axis ax = axis_of(x, a.nx, a.nxm);
    const Axis ay = axis_of(y, a.ny, a.nym);
    // with heights, level_of returns the result of arithmetic; in level-index
    // mode it returns z as it is
    const float lz = level_of(z, lh, a.nz, a.has_heights);
    const Axis az = a.has_heights ? axis_of(lz, a.nz, a.nzm) : axis_of<true>(lz, a.nz, a.nzm);
    const uint32_t ny = a.ny, nz = a.nz;
    const uint32_t r00 = (ax.lo * ny + ay.lo) * nz, r10 = (ax.hi * ny + ay.lo) * nz;
    const uint32_t r01 = (ax.lo * ny + ay.hi) * nz, r11 = (ax.hi * ny + ay.hi) * nz;
    const float ux = 1.0f - ax.t, uy = 1.0f - ay.t, uz = 1.0f - az.t, ua = 1.0f - alpha;
    // One z-plane of corners at a time (x, then y within the plane): the
    // plane values are the x-then-y interpolants of the 8-corner formula, so
    // the result is the same while only 4 corners are held at once.
    float pl[2][6];
#pragma unroll
    for (int zp = 0; zp < 2; ++zp) {
        const uint32_t l = zp ? az.hi : az.lo;
        const uint32_t idx[4] = {r00 + l, r10 + l, r01 + l, r11 + l};
        float c[4][6];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 p = load_at(a.met, idx[k] << 5);
            const float4 q = load_at(a.met, idx[k] << 5, 16u);
            c[k][0] = p.x; c[k][1] = p.y; c[k][2] = p.z;
            c[k][3] = p.w; c[k][4] = q.x; c[k][5] = q.y;
        }
#pragma unroll
        for (int f = 0; f < 6; ++f)
            pl[zp][f] = mix(mix(c[0][f], c[1][f], ax.t, ux), mix(c[2][f], c[3][f], ax.t, ux), ay.t, uy);
    }
    float v[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) v[f] = mix(pl[0][f], pl[1][f], az.t, uz);
    return Wind{mix(v[0], v[3], alpha, ua), mix(v[1], v[4], alpha, ua), mix(v[2], v[5], alpha, ua)};
}

__device__ __forceinline__ void split_cell(float c, uint32_t n, float nm, int32_t &cell, float &frac)
{
    const float cl = clampq(c, 0.0f, nm);
    uint32_t ci = (uint32_t)floorf(cl);
    float fr = cl - (float)ci;
    if (ci >= n - 1u) { ci = n - 1u; fr = 0.0f; }
    cell = (int32_t)ci;
    frac = fr;
}
__device__ __forceinline__ void split_cell(float c, uint32_t n, int32_t &cell, float &frac)
{
    split_cell(c, n, (float)(n - 1), cell, frac);
}

// ---------------------------------------------------------------- height derivatives
//
// The Hanna profiles and the Langevin updates below are templates over the
// scalar type.  The step kernel instantiates them with float.  The adjoint
// instantiates them with Dz, a value carried together with its derivative
// with respect to the particle height: the value part of every Dz operation
// is the float operation the step kernel performs, in the same order, so a
// Dz evaluation reproduces the step's values bit for bit and adds d/dz.

struct Dz { float v, d; };

__device__ __forceinline__ float val(float x) { return x; }
__device__ __forceinline__ float val(Dz x) { return x.v; }
template <class T> __device__ __forceinline__ T kon(float x);
template <> __device__ __forceinline__ float kon<float>(float x) { return x; }
template <> __device__ __forceinline__ Dz kon<Dz>(float x) { return Dz{x, 0.0f}; }

__device__ __forceinline__ Dz operator-(Dz a) { return Dz{-a.v, -a.d}; }
__device__ __forceinline__ Dz operator+(Dz a, Dz b) { return Dz{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dz operator+(Dz a, float b) { return Dz{a.v + b, a.d}; }
__device__ __forceinline__ Dz operator+(float a, Dz b) { return Dz{a + b.v, b.d}; }
__device__ __forceinline__ Dz operator-(Dz a, Dz b) { return Dz{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dz operator-(Dz a, float b) { return Dz{a.v - b, a.d}; }
__device__ __forceinline__ Dz operator-(float a, Dz b) { return Dz{a - b.v, -b.d}; }
__device__ __forceinline__ Dz operator*(Dz a, Dz b) { return Dz{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dz operator*(Dz a, float b) { return Dz{a.v * b, a.d * b}; }
__device__ __forceinline__ Dz operator*(float a, Dz b) { return Dz{a * b.v, a * b.d}; }
__device__ __forceinline__ Dz operator/(Dz a, Dz b)
{
    const float q = a.v / b.v;
    return Dz{q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dz operator/(Dz a, float b) { return Dz{a.v / b, a.d / b}; }
__device__ __forceinline__ Dz operator/(float a, Dz b)
{
    const float q = a / b.v;
    return Dz{q, -q * b.d / b.v};
}

__device__ __forceinline__ float texp(float x) { return fpm_expf(x); }
__device__ __forceinline__ Dz texp(Dz x)
{
    const float e = fpm_expf(x.v);
    return Dz{e, e * x.d};
}
// texp of an argument that cannot be NaN when NOT_NAN holds
template <bool NOT_NAN> __device__ __forceinline__ float texp_nn(float x)
{
    return NOT_NAN ? fpm_expf_not_nan(x) : fpm_expf(x);
}
template <bool NOT_NAN> __device__ __forceinline__ Dz texp_nn(Dz x)
{
    const float e = NOT_NAN ? fpm_expf_not_nan(x.v) : fpm_expf(x.v);
    return Dz{e, e * x.d};
}
__device__ __forceinline__ float tsqrt(float x) { return sqrt_cr(x); }
// (the derivative 1 / (2 sqrt(x)) from the root's own reciprocal square root
// where the short sequence runs)
__device__ __forceinline__ Dz tsqrt(Dz x)
{
    if (__builtin_expect(__ballot(!(x.v >= 0x1p-96f && x.v <= F32_MAX)) != 0ull, 0)) {
        const float r = __builtin_sqrtf(x.v);
        return Dz{r, r > 0.0f ? x.d / (2.0f * r) : 0.0f};
    }
    float hy;
    const float r = sqrt_in_range(x.v, &hy);
    return Dz{r, x.d * hy};
}
// (of an argument known to lie in [2^-96, FLT_MAX]: no range test)
__device__ __forceinline__ float tsqrt_in_range(float x) { return sqrt_in_range(x); }
__device__ __forceinline__ Dz tsqrt_in_range(Dz x)
{
    float hy;
    const float r = sqrt_in_range(x.v, &hy);
    return Dz{r, x.d * hy};
}
// (of an argument known not to be +infinity)
__device__ __forceinline__ float tsqrt_finite(float x) { return sqrt_cr_finite(x); }
__device__ __forceinline__ Dz tsqrt_finite(Dz x)
{
    if (__builtin_expect(__ballot(!(x.v >= 0x1p-96f)) != 0ull, 0)) {
        const float r = __builtin_sqrtf(x.v);
        return Dz{r, r > 0.0f ? x.d / (2.0f * r) : 0.0f};
    }
    float hy;
    const float r = sqrt_in_range(x.v, &hy);
    return Dz{r, x.d * hy};
}
__device__ __forceinline__ float tabs(float x) { return fabsf(x); }
__device__ __forceinline__ Dz tabs(Dz x) { return Dz{fabsf(x.v), x.v < 0.0f ? -x.d : x.d}; }
__device__ __forceinline__ float tmax(float x, float lo) { return fmaxf(x, lo); }
__device__ __forceinline__ Dz tmax(Dz x, float lo)
{
    const float r = fmaxf(x.v, lo);
    return Dz{r, r == x.v ? x.d : 0.0f};
}
// (the profiles clamp heights and height ratios produced by arithmetic)
__device__ __forceinline__ float tclamp(float x, float lo, float hi) { return clampq(x, lo, hi); }
__device__ __forceinline__ Dz tclamp(Dz x, float lo, float hi)
{
    return Dz{clampq(x.v, lo, hi), (x.v > lo && x.v < hi) ? x.d : 0.0f};
}
__device__ __forceinline__ float tpos(float x) { return pos_or_zero(x); }
__device__ __forceinline__ Dz tpos(Dz x) { return (finite_f(x.v) && x.v > 0.0f) ? x : Dz{0.0f, 0.0f}; }
// Quotients of the Hanna profiles and the Langevin updates.  LEAN: the
// operands are known to lie in the range where fpm_div_unscaled equals the
// IEEE division (TurbRegime::lean); the height-derivative type keeps the
// IEEE division, whose value part is then the same.
template <bool LEAN> __device__ __forceinline__ float qdiv(float n, float d)
{
    return LEAN ? fpm_div_unscaled(n, d) : n / d;
}
// (Under LEAN the height derivative multiplies by the refined reciprocal of
// the denominator instead of dividing by it: the value is the same as the
// forward step's, the derivative within a unit or two in the last place.)
template <bool LEAN> __device__ __forceinline__ Dz qdiv(Dz n, Dz d)
{
    if (!LEAN) return n / d;
    float r;
    const float q = fpm_div_unscaled_rcp(n.v, d.v, &r);
    return Dz{q, (n.d - q * d.d) * r};
}
template <bool LEAN> __device__ __forceinline__ Dz qdiv(Dz n, float d)
{
    if (!LEAN) return n / d;
    float r;
    const float q = fpm_div_unscaled_rcp(n.v, d, &r);
    return Dz{q, n.d * r};
}
template <bool LEAN> __device__ __forceinline__ Dz qdiv(float n, Dz d)
{
    if (!LEAN) return n / d;
    float r;
    const float q = fpm_div_unscaled_rcp(n, d.v, &r);
    return Dz{q, -q * d.d * r};
}

// The binary64 log/exp tables of fp_math.h, copied into LDS by the kernels
// that evaluate the Hanna powers (load_math_tab, before the kernel's first
// barrier): the same values as the device constants, read at LDS latency
// rather than through the vector cache that the state and met traffic stream
// through.  A kernel that calls the lds_* functions must load the copy first.
constexpr uint32_t kMathTab = 256 + 32;
__device__ __forceinline__ double *math_tab()
{
    __shared__ double t[kMathTab];
    return t;
}
template <uint32_t BLOCK>
__device__ __forceinline__ void load_math_tab()
{
    double *t = math_tab();
    for (uint32_t k = threadIdx.x; k < kMathTab; k += BLOCK) t[k] = k < 256u ? FPM_LOG_TAB[k] : FPM_EXP2_J[k - 256u];
}
__device__ __forceinline__ double lds_log_d(double x) { return fpm_log_d_tab(x, math_tab()); }
__device__ __forceinline__ float lds_exp_d_to_f(double t) { return fpm_exp_d_to_f_tab(t, math_tab() + 256); }
__device__ __forceinline__ float lds_powf(float x, float y)
{
    return fpm_powf_tab(x, y, math_tab(), math_tab() + 256);
}

// zeta^y as exp(y log zeta) in binary64 with log zeta = L precomputed; the
// _in_range forms for y L known to lie in [-104, 89]
__device__ __forceinline__ float pow_log_in_range(float, float y, double L)
{
    return fpm_exp_d_to_f_in_range((double)y * L, math_tab() + 256);
}
__device__ __forceinline__ float pow_log_floor_in_range(float, float y, bool use, double L, double Lfloor)
{
    return fpm_exp_d_to_f_in_range((double)y * (use ? L : Lfloor), math_tab() + 256);
}
__device__ __forceinline__ Dz pow_log_in_range(Dz zeta, float y, double L)
{
    const float v = fpm_exp_d_to_f_in_range((double)y * L, math_tab() + 256);
    return Dz{v, zeta.v > 0.0f ? v * y / zeta.v * zeta.d : 0.0f};
}
__device__ __forceinline__ Dz pow_log_floor_in_range(Dz zeta, float y, bool use, double L, double Lfloor)
{
    const float v = fpm_exp_d_to_f_in_range((double)y * (use ? L : Lfloor), math_tab() + 256);
    return Dz{v, (use && zeta.v > 0.0f) ? v * y / zeta.v * zeta.d : 0.0f};
}
__device__ __forceinline__ float pow_log(float, float y, double L) { return lds_exp_d_to_f((double)y * L); }
__device__ __forceinline__ Dz pow_log(Dz zeta, float y, double L)
{
    const float v = lds_exp_d_to_f((double)y * L);
    return Dz{v, zeta.v > 0.0f ? v * y / zeta.v * zeta.d : 0.0f};
}
// max(zeta, floor)^y with log of the floor Lfloor: exp(y (use ? L : Lfloor))
__device__ __forceinline__ float pow_log_floor(float, float y, bool use, double L, double Lfloor)
{
    return lds_exp_d_to_f((double)y * (use ? L : Lfloor));
}
__device__ __forceinline__ Dz pow_log_floor(Dz zeta, float y, bool use, double L, double Lfloor)
{
    const float v = lds_exp_d_to_f((double)y * (use ? L : Lfloor));
    return Dz{v, (use && zeta.v > 0.0f) ? v * y / zeta.v * zeta.d : 0.0f};
}

// ---------------------------------------------------------------- Hanna

// The turbulence regime of a particle is fixed for the whole step: it depends
// on L and the mixing height of its PBL cell only (langevin_fused.wgsl:236).
struct TurbRegime {
    float ust, wst, ol, h;
    float sigu_cell;  // the cell's unstable sigma_u (second PBL record)
    bool neu, uns;    // neutral, unstable (else stable)
    // wavefront-uniform: every lane's cell and height are in the range where
    // each quotient of the profiles and of the Langevin updates needs no
    // range scaling (lean_lane)
    bool lean;
    // wavefront-uniform: with lean, the regime every active lane of the Hanna
    // evaluation shares (0 neutral, 1 unstable, 2 stable) or 3 for lanes in
    // different regimes; -1 without lean.  It holds for the later profile
    // evaluations of the step, whose active lanes are the same.
    int wr;
};
template <class T> struct TurbT : TurbRegime {
    T sigu, sigv, sigw, dsigwdz, tlu, tlv, tlw;
};
using Turb = TurbT<float>;

__device__ __forceinline__ bool neutral_regime(float ol, float h)
{
    return !finite_f(ol) || ol == 0.0f || h / fabsf(ol) < 1.0f;
}

// Raw vertical profile at (z, zeta) before the floors: the w part of
// compute_hanna (langevin_fused.wgsl:224-293) and hanna_short_w (:296-344),
// which share their formulas.
//   neutral : corr = z/u*, sigma_w = 1.3 u* e^{-2e-4 corr} (+0.01),
//             T_Lw = 0.5 z / sigma_w / (1 + 1.5e-3 corr)
//   unstable: sigma_w^2 = 1.2 w*^2 (1-0.9 zeta) zeta^(2/3) + (1.8-1.4 zeta) u*^2 ...
//   stable  : sigma_w = 0.01 + 1.3 u* (1-zeta), T_Lw = 0.1 h / sigma_w zeta^0.8
// Slots shared by all regimes: d0 (one division), E (one exp), d1 and d2 (two
// divisions by sigma_w and by h or 1+1.5e-3 corr).  zeta^(2/3), zeta^0.8 and
// max(zeta,1e-3)^(-1/3) are exp(y log zeta) in binary64 and share one log
// (fpm_powf's own evaluation, so the values are those of fpm_powf).
template <class T> struct VRawT { T sigw, dsigwdz, tlw, d0, d2; };

// R: the regime of every active lane of the wavefront (0 neutral, 1 unstable,
// 2 stable), or -1 for a wavefront that mixes regimes (selected per lane).
// The arithmetic of each regime is the same in every instantiation.
template <int R, bool LEAN, class T>
__device__ __forceinline__ VRawT<T> vertical_raw_r(const TurbRegime &t, T z, T zeta, double log_1em3)
{
    const float ust = t.ust, wst = t.wst, ol = t.ol, h = t.h;
    const bool neu = R < 0 ? t.neu : R == 0;
    const bool uns = R < 0 ? t.uns : R == 1;
    VRawT<T> v;
    // d0: neutral z/u* (corr), unstable z/L, stable -1.3 u* / h
    v.d0 = qdiv<LEAN>(uns || neu ? z : kon<T>(-1.3f * ust), neu ? ust : (uns ? ol : h));
    // E: neutral e^{-2e-4 corr}, unstable e^{-5 zeta}; d2 below serves the
    // same two regimes, so a wavefront of stable particles skips both
    const bool any_nu = R < 0 ? __ballot(neu || uns) != 0ull : R != 2;
    T E = kon<T>(0.0f);
    // (under LEAN d0 and zeta are finite)
    if (any_nu) E = texp_nn<LEAN>(neu ? -2.0e-4f * v.d0 : -5.0f * zeta);
    T pa = kon<T>(0.0f), pb = kon<T>(0.0f);
    if (!neu && LEAN) {
        // zeta in [2^-60, 1] (lean_lane: z >= 2^-40, h <= 2^20), so log zeta
        // lies in [-41.6, 0], the exponents 0.8, 2/3 and -1/3 times it in
        // [-33.3, 13.9] (log 1e-3 = -6.9 too): the zeta > 0 tests are true and
        // the exponentials need neither clamp nor overflow/underflow selects
        const double L = lds_log_d((double)val(zeta));
        pa = pow_log_in_range(zeta, uns ? 2.0f / 3.0f : 0.8f, L);
        if (uns) pb = pow_log_floor_in_range(zeta, -1.0f / 3.0f, val(zeta) >= 1.0e-3f, L, log_1em3);
    } else if (!neu) {
        const double L = lds_log_d((double)(val(zeta) > 0.0f ? val(zeta) : 1.0f));
        pa = val(zeta) > 0.0f ? pow_log(zeta, uns ? 2.0f / 3.0f : 0.8f, L) : kon<T>(0.0f);
        if (uns) pb = pow_log_floor(zeta, -1.0f / 3.0f, val(zeta) >= 1.0e-3f, L, log_1em3);
    }
    T sigw, sraw = kon<T>(0.0f);
    if (neu) {
        sraw = 1.3f * ust * E;
        sigw = sraw + SIGMA_FLOOR;
    } else if (uns) {
        const T sq = 1.2f * wst * wst * (1.0f - 0.9f * zeta) * pa + (1.8f - 1.4f * zeta) * ust * ust;
        // under LEAN sq lies in [0.4 u*^2, 2^43] (u* >= 1e-4 and u*, w* <= 2^20)
        sigw = (LEAN ? tsqrt_in_range(sq) : tsqrt(sq)) + SIGMA_FLOOR;
    } else {
        sigw = SIGMA_FLOOR + 1.3f * ust * (1.0f - zeta);
    }
    // d1: neutral 0.5 z / sigma_w, unstable 0.5 / sigma_w, stable 0.1 h / sigma_w
    const T d1 = qdiv<LEAN>(neu ? 0.5f * z : kon<T>(uns ? 0.5f : 0.1f * h), sigw);
    // d2: neutral d1 / (1 + 1.5e-3 corr), unstable d1 / h
    v.d2 = kon<T>(0.0f);
    if (any_nu) v.d2 = qdiv<LEAN>(d1, neu ? 1.0f + 1.5e-3f * v.d0 : kon<T>(h));
    if (neu) {
        v.dsigwdz = -2.0e-4f * sraw;
        v.tlw = v.d2;
    } else if (uns) {
        v.dsigwdz = v.d2 * (-1.4f * ust * ust + wst * wst * (0.8f * pb - 1.8f * pa));
        const bool ca = val(z) < fabsf(ol);
        const bool cb = !ca && val(zeta) < 0.1f;
        const T d3 = qdiv<LEAN>(ca ? 0.1f * z : (cb ? 0.59f * z : kon<T>(0.15f * h)),
                                ca ? sigw * (0.55f - 0.38f * tabs(v.d0)) : sigw);
        v.tlw = (ca || cb) ? d3 : d3 * (1.0f - E);
    } else {
        v.dsigwdz = v.d0;
        v.tlw = d1 * pa;
    }
    v.sigw = sigw;
    return v;
}

// hanna_short_w (langevin_fused.wgsl:296-344) for one regime R of every lane
// (or R = -1, selected per lane), with the floors of :335-339
template <int R, bool LEAN, class T>
__device__ __forceinline__ void vertical_profile_r(const TurbRegime &t, T z_in, double log_1em3, T &sigw,
                                                   T &dsigwdz, T &tlw)
{
    const T z = tclamp(z_in, 0.0f, t.h);
    const T zeta = tclamp(qdiv<LEAN>(z, t.h), 0.0f, 1.0f);
    const VRawT<T> v = vertical_raw_r<R, LEAN>(t, z, zeta, log_1em3);
    tlw = tmax(v.tlw, TLW_MIN);
    dsigwdz = val(v.dsigwdz) == 0.0f ? kon<T>(1.0e-10f) : v.dsigwdz;
    sigw = v.sigw;
    // Under LEAN sigma_w is finite and >= 0.01, T_L_w and dsigma_w/dz finite
    // (lean_lane), so these floors change nothing
    if (!LEAN) {
        if (!finite_f(val(v.sigw)) || val(v.sigw) <= 0.0f) sigw = kon<T>(SIGMA_FLOOR);
        if (!finite_f(val(tlw))) tlw = kon<T>(TLW_MIN);
        if (!finite_f(val(dsigwdz))) dsigwdz = kon<T>(1.0e-10f);
    }
}

// The regime-selected evaluation: a wavefront whose active lanes pass the
// range check (t.wr >= 0, decided once per step by hanna_full) runs with
// unscaled quotients, in its regime's instantiation when its lanes share one
// (no per-lane selects between the regimes' operands) or in the per-lane
// selected one; any other wavefront runs the per-lane selected one with IEEE
// quotients.  The arithmetic of each regime is the same in every
// instantiation, and under t.lean both quotient forms give the same values.
template <class T>
__device__ __forceinline__ void vertical_profile(const TurbRegime &t, T z_in, double log_1em3, T &sigw,
                                                 T &dsigwdz, T &tlw)
{
    if (t.wr == 0) return vertical_profile_r<0, true>(t, z_in, log_1em3, sigw, dsigwdz, tlw);
    if (t.wr == 1) return vertical_profile_r<1, true>(t, z_in, log_1em3, sigw, dsigwdz, tlw);
    if (t.wr == 2) return vertical_profile_r<2, true>(t, z_in, log_1em3, sigw, dsigwdz, tlw);
    if (t.wr == 3) return vertical_profile_r<-1, true>(t, z_in, log_1em3, sigw, dsigwdz, tlw);
    vertical_profile_r<-1, false>(t, z_in, log_1em3, sigw, dsigwdz, tlw);
}

// The per-cell part of compute_hanna, evaluated once per PBL upload: the
// Obukhov length (obukhov_length_from_inverse, langevin_fused.wgsl:213-218),
// the regime (0 neutral, 1 unstable, 2 stable) and the unstable regime's
// height-independent sigma_u.  Stored beside {u*, w*, hmix, 1/L} as the
// cell's second 16-byte record.
__device__ __forceinline__ float4 pbl_cell_terms_dev(float4 pb)
{
    const float ust = fmaxf(pb.x, USTAR_MIN), h = fmaxf(pb.z, H_MIN);
    const float ol = fabsf(pb.w) < 1.0e-5f ? __builtin_inff() : 1.0f / pb.w;
    const bool neu = neutral_regime(ol, h), uns = !neu && ol < 0.0f;
    const float sigu = uns ? SIGMA_FLOOR + ust * fpm_powf(12.0f - 0.5f * h / ol, 1.0f / 3.0f) : 0.0f;
    return make_float4(ol, sigu, __uint_as_float(neu ? 0u : (uns ? 1u : 2u)), 0.0f);
}

// compute_hanna (langevin_fused.wgsl:224-293) from the cell's two records,
// for a wavefront in one regime R or a mixed one (R = -1), as vertical_raw_r
template <int R, bool LEAN, class T>
__device__ __forceinline__ TurbT<T> hanna_full_r(TurbT<T> t, T z_in, double log_1em3)
{
    const float ust = t.ust, h = t.h;
    const bool neu = R < 0 ? t.neu : R == 0;
    const bool uns = R < 0 ? t.uns : R == 1;
    const T z = tclamp(z_in, 0.0f, h);
    const T zeta = tclamp(qdiv<LEAN>(z, h), 0.0f, 1.0f);
    const VRawT<T> v = vertical_raw_r<R, LEAN>(t, z, zeta, log_1em3);
    T sigu, sigv, tlu, tlv;
    if (neu) {
        sigu = SIGMA_FLOOR + 2.0f * ust * texp_nn<LEAN>(-3.0e-4f * v.d0);
        sigv = v.sigw;
    } else if (uns) {
        sigu = kon<T>(t.sigu_cell);
        sigv = sigu;
    } else {
        sigu = SIGMA_FLOOR + 2.0f * ust * (1.0f - zeta);
        sigv = SIGMA_FLOOR + 1.3f * ust * (1.0f - zeta);
    }
    // unstable and stable: 0.15 h / sigma_u; neutral: T_Lu = T_Lw
    T d5 = kon<T>(0.0f);
    if (R < 0 ? __ballot(!neu) != 0ull : R != 0) d5 = qdiv<LEAN>(0.15f * h, sigu);
    if (neu) {
        tlu = v.d2;
        tlv = tlu;
    } else if (uns) {
        tlu = d5;
        tlv = tlu;
    } else {
        // zeta in [0, 1]; under LEAN zeta = z / h >= 2^-60
        tlu = d5 * (LEAN ? tsqrt_in_range(zeta) : tsqrt_finite(zeta));
        tlv = 0.467f * tlu;
    }
    T sigw = v.sigw, dsigwdz = v.dsigwdz, tlw = v.tlw;
    tlu = tmax(tlu, TLU_MIN);
    tlv = tmax(tlv, TLV_MIN);
    tlw = tmax(tlw, TLW_MIN);
    if (val(dsigwdz) == 0.0f) dsigwdz = kon<T>(1.0e-10f);
    // Under LEAN every sigma is finite and >= 0.01 (the cell's unstable
    // sigma_u too: 0.01 + u* (12 + 0.5 h/|L|)^(1/3) <= 2^34) and every T_L and
    // dsigma_w/dz finite (lean_lane), so these floors change nothing
    if (!LEAN) {
        if (!finite_f(val(sigu)) || val(sigu) <= 0.0f) sigu = kon<T>(SIGMA_FLOOR);
        if (!finite_f(val(sigv)) || val(sigv) <= 0.0f) sigv = kon<T>(SIGMA_FLOOR);
        if (!finite_f(val(sigw)) || val(sigw) <= 0.0f) sigw = kon<T>(SIGMA_FLOOR);
        if (!finite_f(val(tlu))) tlu = kon<T>(TLU_MIN);
        if (!finite_f(val(tlv))) tlv = kon<T>(TLV_MIN);
        if (!finite_f(val(tlw))) tlw = kon<T>(TLW_MIN);
        if (!finite_f(val(dsigwdz))) dsigwdz = kon<T>(1.0e-10f);
    }
    t.sigu = sigu; t.sigv = sigv; t.sigw = sigw; t.dsigwdz = dsigwdz;
    t.tlu = tlu; t.tlv = tlv; t.tlw = tlw;
    return t;
}

// The lane's range check for unscaled quotients (fpm_div_unscaled): u*, w*
// and h at most 2^20, the height at least 2^-40 and, in the unstable regime,
// |L| at least 2^-20; the launch adds dt in [2^-20, 2^20] and z_min in
// [2^-40, 1) (so that every sub-step height lies in [z_min, h]).  Then every
// numerator and denominator of compute_hanna, hanna_short_w and the two
// Ornstein-Uhlenbeck updates lies in [2^-47, 2^41] and every quotient in
// [2^-90, 2^40]: sigma_w, sigma_u >= 0.01 (the 0.01 terms, the unstable
// sigma_w^2 >= 0.4 u*^2), all sigmas <= 2^34, 0.55 - 0.38 z/|L| > 0.17 where
// it divides, 1 + 1.5e-3 z/u* <= 2^24, T_L in [10, 2^27].
__device__ __forceinline__ bool lean_lane(float ust, float wst, float h, float z, bool uns, float ol)
{
    return fmaxf(fmaxf(ust, wst), h) <= 0x1p20f && z >= 0x1p-40f && (!uns || fabsf(ol) >= 0x1p-20f);
}

// launch_lean: the launch's part of the range check (StepArgs::lean_div)
template <class T>
__device__ __forceinline__ TurbT<T> hanna_full(float4 pb, float4 pc, T z_in, double log_1em3, bool launch_lean = false)
{
    TurbT<T> t;
    t.ust = fmaxf(pb.x, USTAR_MIN);
    t.wst = fmaxf(pb.y, 0.0f);
    t.h = fmaxf(pb.z, H_MIN);
    t.ol = pc.x;
    t.sigu_cell = pc.y;
    const uint32_t regime = __float_as_uint(pc.z);
    t.neu = regime == 0u;
    t.uns = regime == 1u;
    t.lean = launch_lean && __ballot(!lean_lane(t.ust, t.wst, t.h, val(z_in), t.uns, t.ol)) == 0ull;
    t.wr = -1;
    if (t.lean) {
        const uint64_t live = __ballot(1), n = __ballot(t.neu), u = __ballot(t.uns);
        t.wr = n == live ? 0 : (u == live ? 1 : ((n | u) == 0ull ? 2 : 3));
    }
    if (t.wr == 0) return hanna_full_r<0, true>(t, z_in, log_1em3);
    if (t.wr == 1) return hanna_full_r<1, true>(t, z_in, log_1em3);
    if (t.wr == 2) return hanna_full_r<2, true>(t, z_in, log_1em3);
    if (t.wr == 3) return hanna_full_r<-1, true>(t, z_in, log_1em3);
    return hanna_full_r<-1, false>(t, z_in, log_1em3);
}

// ---------------------------------------------------------------- Langevin
//
// update_horizontal_component / update_vertical_component
// (langevin_fused.wgsl:350-389).  Both branches of the reference,
//   r <  0.5: (1-r) u + eta sigma sqrt(2r)          [+ dt drift]
//   r >= 0.5: e^{-r} u + eta sigma sqrt(1-e^{-2r})  [+ T_L (1-e^{-r}) drift]
// are one expression with selected coefficients: one division, one exp and
// one sqrt per update whichever branch a lane takes.
//
// Precondition: sigma and T_L come from compute_hanna / hanna_short_w after
// their floors (sigma finite and > 0, T_L finite and >= its 10 s or 30 s
// floor).  The reference's clamps of both to [0, inf) and its T_L == 0 case
// are then the identity and false, and are not evaluated.

// BR: the branch of every lane of the wavefront (0: r < 0.5, 1: r >= 0.5) or
// -1 for a wavefront with both (selected per lane).  Under LEAN, r lies in
// [2^-47, 2^41] (dt >= 2^-20, T_L in [10, 2^27]) and e^{-r} in [0, 1], so the
// reference's clamps of 2 r and 1 - e^{-2r} to [0, inf) are the identity and
// the root's argument lies in [2^-46, 1] (2 r >= 2^-46 where r < 0.5,
// 1 - e^{-2r} >= 1 - e^{-1} where r >= 0.5): no range test.
template <int BR, bool LEAN, class T>
__device__ __forceinline__ T ou_horizontal_b(T prev, T sigma, T r, float eta, float *coef)
{
    const bool small = BR < 0 ? val(r) < 0.5f : BR == 0;
    // e^{-r} serves the r >= 0.5 branch only: skipped by a wavefront without one
    T e = kon<T>(0.0f);
    if (BR < 0 ? __ballot(!small) != 0ull : BR == 1) e = texp_nn<LEAN>(-r);
    const T arg = small ? 2.0f * r : 1.0f - e * e;
    const T root = LEAN ? tsqrt_in_range(arg) : tsqrt_finite(tmax(arg, 0.0f));
    const T c = small ? 1.0f - r : e;
    if (coef) *coef = val(c);
    const T nx = c * prev + eta * sigma * root;
    return !finite_f(val(nx)) ? kon<T>(0.0f) : nx;
}

template <int BR, bool LEAN, class T>
__device__ __forceinline__ T ou_vertical_b(T prev, T sw, T dsigwdz, T tl, T r, float dt, float rho, float eta,
                                           float *coef)
{
    const T drift = sw * (dsigwdz + rho * sw);
    const bool small = BR < 0 ? val(r) < 0.5f : BR == 0;
    T e = kon<T>(0.0f);
    if (BR < 0 ? __ballot(!small) != 0ull : BR == 1) e = texp_nn<LEAN>(-r);
    const T arg = small ? 2.0f * r : 1.0f - e * e;
    const T root = LEAN ? tsqrt_in_range(arg) : tsqrt_finite(tmax(arg, 0.0f));
    const T c = small ? 1.0f - r : e;
    if (coef) *coef = val(c);
    const T nx = c * prev + eta * sw * root + (small ? dt * drift : tl * (1.0f - e) * drift);
    return !finite_f(val(nx)) ? kon<T>(0.0f) : nx;
}

// A wavefront whose lanes all take one branch runs that branch alone (under
// LEAN; the arithmetic of each branch is the same as in the per-lane form).
// coef, if given, receives the retained fraction of the old velocity, 1 - r
// or e^{-r} (d new / d prev at fixed sigma and T_L).
template <bool LEAN = false, class T>
__device__ __forceinline__ T ou_horizontal(T prev, T sigma, T tl, float dt, float eta, float *coef = nullptr)
{
    const T r = qdiv<LEAN>(dt, tl);
    if (LEAN) {
        const uint64_t sm = __ballot(val(r) < 0.5f), live = __ballot(1);
        if (sm == live) return ou_horizontal_b<0, LEAN>(prev, sigma, r, eta, coef);
        if (sm == 0ull) return ou_horizontal_b<1, LEAN>(prev, sigma, r, eta, coef);
    }
    return ou_horizontal_b<-1, LEAN>(prev, sigma, r, eta, coef);
}

template <bool LEAN = false, class T>
__device__ __forceinline__ T ou_vertical(T prev, T sw, T dsigwdz, T tl, float dt, float rho, float eta,
                                         float *coef = nullptr)
{
    const T r = qdiv<LEAN>(dt, tl);
    if (LEAN) {
        const uint64_t sm = __ballot(val(r) < 0.5f), live = __ballot(1);
        if (sm == live) return ou_vertical_b<0, LEAN>(prev, sw, dsigwdz, tl, r, dt, rho, eta, coef);
        if (sm == 0ull) return ou_vertical_b<1, LEAN>(prev, sw, dsigwdz, tl, r, dt, rho, eta, coef);
    }
    return ou_vertical_b<-1, LEAN>(prev, sw, dsigwdz, tl, r, dt, rho, eta, coef);
}

// PBL reflection (langevin_fused.wgsl:481-500)
__device__ __forceinline__ void reflect(float &pz, float &tw, int32_t &cbt, float zmin, float h)
{
    // Branch-free form of the reference loop: every lane evaluates each bounce
    // with selects and the wavefront leaves when no lane has a bounce left.  A
    // lane inside [z_min, h] (or NaN) is left as it is, which is where the
    // reference's loop breaks; a lane with h <= z_min is not touched, its early
    // return.
    const bool on = h > zmin;
    const float lo2 = 2.0f * zmin, hi2 = 2.0f * h;
    float z = pz;
    for (int bounce = 0; bounce < 10; ++bounce) {
        // with h > z_min, z > h excludes z < z_min (the reference's else-if)
        const bool below = on && z < zmin;
        const bool above = on && z > h;
        if (__ballot(below || above) == 0ull) break;
        z = below ? lo2 - z : (above ? hi2 - z : z);
        const bool flip = below || above;
        cbt = flip ? -cbt : cbt;
        tw = flip ? -tw : tw;
    }
    if (on) pz = clampq(z, zmin, h);
}

// ---------------------------------------------------------------- gridding

// FAST (g.fast, for the fused step's cells: cx and cy index the met grid):
// cx * refine_x + fx and cy * refine_y + fy are non-negative and below 2^31,
// so 32-bit arithmetic gives the same cell and only the upper clamp can act;
// with non-decreasing output heights the levels below z are a prefix, and the
// first k with z <= h[k] is their count (nz - 1 when there is none, NaN z
// included).
template <bool FAST = false>
__device__ __forceinline__ uint32_t grid_cell(const GridArgs &g, const float *oh, int32_t cx, int32_t cy,
                                              float px, float py, float z)
{
    int32_t fx = (int32_t)floorf(px * (float)g.refine_x);
    int32_t fy = (int32_t)floorf(py * (float)g.refine_y);
    fx = max(0, min(fx, (int32_t)g.refine_x - 1));
    fy = max(0, min(fy, (int32_t)g.refine_y - 1));
    if (FAST) {
        const uint32_t ix = min((uint32_t)cx * g.refine_x + (uint32_t)fx, g.nx - 1u);
        const uint32_t iy = min((uint32_t)cy * g.refine_y + (uint32_t)fy, g.ny - 1u);
        uint32_t iz;
        if (g.metres) {
            uint32_t below = 0;
            for (uint32_t k = 0; k < g.nz; ++k) below += !(z <= oh[k]) ? 1u : 0u;
            iz = min(below, g.nz - 1u);
        } else {
            const int32_t l = (int32_t)floorf(z);
            iz = (uint32_t)max(0, min(l, (int32_t)g.nz - 1));
        }
        return (ix * g.ny + iy) * g.nz + iz;
    }
    int64_t ix = (int64_t)cx * g.refine_x + fx;
    int64_t iy = (int64_t)cy * g.refine_y + fy;
    ix = ix < 0 ? 0 : (ix > (int64_t)g.nx - 1 ? (int64_t)g.nx - 1 : ix);
    iy = iy < 0 ? 0 : (iy > (int64_t)g.ny - 1 ? (int64_t)g.ny - 1 : iy);
    uint32_t iz;
    if (g.metres) {
        // first level with z <= outheights[k], else the top (concentration_gridding.wgsl:81-93)
        iz = g.nz - 1u;
        for (uint32_t k = g.nz; k-- > 0;)
            if (z <= oh[k]) iz = k;
    } else {
        int32_t l = (int32_t)floorf(z);
        iz = (uint32_t)max(0, min(l, (int32_t)g.nz - 1));
    }
    return ((uint32_t)ix * g.ny + (uint32_t)iy) * g.nz + iz;
}

__device__ __forceinline__ uint32_t quantize_mass(float m_in, float scale)
{
    const float m = fmaxf(m_in, 0.0f);
    if (m <= 0.0f) return 0u;
    const float s = m * scale + 0.5f;
    return s >= 4294967295.0f ? 0xffffffffu : (uint32_t)s;
}

}  // namespace

// Sum of v over the wave, complete in lane 63 (the other lanes hold partial
// prefix sums).  DPP row shifts build the 16-lane row sums and the two row
// broadcasts carry them across rows: six VALU operations with no LDS traffic.
__device__ __forceinline__ uint32_t wave_sum_to_lane63(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false); // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false); // row_bcast:31
    return v;
}

// Block-private concentration grid in LDS: an open-addressing hash table of
// (cell, count, mass) that absorbs a block's contributions; the block adds
// each occupied slot to the global grid once at its end.  Point releases put
// most particles of a launch into a handful of output cells, and without the
// table every wave would issue atomics to those same addresses, which the
// memory system serialises.  A cell that finds no free slot within
// kGridProbes probes goes straight to the global grid.  All sums are integer
// sums modulo 2^32, so neither the table nor the order of the adds changes
// the result.
constexpr uint32_t kGridSlotBits = 9;
constexpr uint32_t kGridSlots = 1u << kGridSlotBits;
constexpr uint32_t kGridProbes = 8;
constexpr uint32_t kGridEmpty = 0xffffffffu;  // no cell index reaches 2^32 - 1

struct GridLds {
    uint32_t key[kGridSlots];
    uint32_t cnt[kGridSlots];
    uint32_t mass[kGridSlots];
};

// Empties the table; the caller's barrier must follow before first use.
template <uint32_t BLOCK>
__device__ __forceinline__ void grid_lds_clear(GridLds &t)
{
    for (uint32_t k = threadIdx.x; k < kGridSlots; k += BLOCK) {
        t.key[k] = kGridEmpty;
        t.cnt[k] = 0u;
        t.mass[k] = 0u;
    }
}

// Workgroup-scope atomics for the table: besides stating the scope, they
// keep the compiler from merging a table add and a global-grid add into one
// flat atomic with a selected address.
__device__ __forceinline__ void lds_add(uint32_t *p, uint32_t v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t lds_cas(uint32_t *p, uint32_t expected, uint32_t desired)
{
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    return expected;
}

__device__ __forceinline__ void grid_lds_add(GridLds &t, const GridArgs &g, uint32_t cell, uint32_t c, uint32_t m)
{
    uint32_t h = (cell * 0x9E3779B1u) >> (32u - kGridSlotBits);
    for (uint32_t p = 0; p < kGridProbes; ++p) {
        const uint32_t prev = lds_cas(&t.key[h], kGridEmpty, cell);
        if (prev == kGridEmpty || prev == cell) {
            lds_add(&t.cnt[h], c);
            lds_add(&t.mass[h], m);
            return;
        }
        h = (h + 1u) & (kGridSlots - 1u);
    }
    atomicAdd(&g.counts[cell], c);
    if (m) atomicAdd(&g.mass_fp[cell], m);
}

// Adds the table to the global grid; every thread of the block must call it.
template <uint32_t BLOCK>
__device__ __forceinline__ void grid_lds_flush(GridLds &t, const GridArgs &g)
{
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < kGridSlots; k += BLOCK) {
        const uint32_t cell = t.key[k];
        if (cell == kGridEmpty) continue;
        atomicAdd(&g.counts[cell], t.cnt[k]);
        const uint32_t m = t.mass[k];
        if (m) atomicAdd(&g.mass_fp[cell], m);
    }
}

// Wave-level pre-reduction into the block's table.  Every lane of the wave
// must call this (lanes without a contribution pass valid = false).  The
// lowest contributing lane's cell is read with v_readlane, the lanes with
// that cell are counted and their masses summed with a DPP reduction, and
// that total enters the table once; every other contribution enters on its
// own, in parallel (the table is the block-level reduction).  One pass takes
// a plume's hot cell out of the table's atomics; iterating the passes over
// every distinct cell of a spread-out wave cost more instructions than the
// table inserts they saved.
__device__ __forceinline__ void grid_reduce_wave(GridLds &t, const GridArgs &g, bool valid, uint32_t cell,
                                                 uint32_t q)
{
    const uint64_t pending = __ballot(valid);
    if (!pending) return;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__builtin_ctzll(pending);
    const uint32_t lc = (uint32_t)__builtin_amdgcn_readlane((int)cell, (int)leader);
    const bool mine = valid && cell == lc;
    const uint64_t same = __ballot(mine);
    const uint32_t v = wave_sum_to_lane63(mine ? q : 0u);
    const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)v, (int)(kWave - 1u));
    const bool lead = lane == leader;
    if (lead || (valid && !mine))
        grid_lds_add(t, g, lead ? lc : cell, lead ? (uint32_t)__popcll(same) : 1u, lead ? tot : q);
}

// XCD-aware chunking: the hardware deals blocks round-robin over the 8
// XCDs; give each XCD a contiguous run of particle chunks so that
// Morton-neighbouring particles share one L2.
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t b, uint32_t nb)
{
    const uint32_t xcd = b % kXcds, k = b / kXcds, per = nb / kXcds, rem = nb % kXcds;
    return xcd < rem ? xcd * (per + 1u) + k : rem * (per + 1u) + (xcd - rem) * per + k;
}

// xcd_chunk within each of `rounds` consecutive parts of the grid (each a
// multiple of 8 blocks but the last): every XCD takes a contiguous run from
// every part, so it works on `rounds` regions of the particle range instead
// of one.  Particles of one region differ in cost (their plume's turbulence
// regimes and reflections), and an XCD whose single region is expensive
// finishes up to 25 % after the others; a few regions per XCD even that out
// while each run still shares its met gathers in one L2.  (Finer interleaving
// loses the L2 reuse: dealing blocks round-robin is 40-80 % slower.)
__device__ __forceinline__ uint32_t xcd_chunk_rounds(uint32_t b, uint32_t nb, uint32_t rounds)
{
    const uint32_t part = (nb + rounds * kXcds - 1u) / (rounds * kXcds) * kXcds;
    const uint32_t r0 = b / part * part;
    return r0 + xcd_chunk(b - r0, min(part, nb - r0));
}


}  // namespace fp
