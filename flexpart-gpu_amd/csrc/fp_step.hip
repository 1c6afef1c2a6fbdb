// fp_step.hip - the fused per-particle timestep for gfx950.
//
// One kernel replaces the reference's four production dispatches
// (advection_dual_wind.wgsl, langevin_fused.wgsl, dry_deposition.wgsl,
// wet_deposition.wgsl; step order of particle_step.wgsl:684-883).  Each
// particle's hot state (44 B) is read once from SoA arrays with coalesced
// loads, advanced through advection, Hanna turbulence, Philox/Box-Muller
// noise, the Langevin update with vertical sub-steps and PBL reflection, dry
// and wet deposition, and written back once (56 B with the velocity memory).
// Optionally the concentration grid is accumulated in the same pass.
//
// The kernel is bound by vector-ALU issue, not by HBM: its arithmetic is laid
// out so that a wavefront whose particles sit in different turbulence
// regimes (neutral, unstable, stable Hanna profiles) or Langevin branches
// still evaluates each expensive primitive - division, exp, the binary64
// log/exp behind pow, sqrt - once, with regime-selected operands, instead of
// once per divergent branch.  Every selected operation is the same operation
// on the same operands as in the reference formula, so values are unchanged.
//
// Floating point: compiled with -ffp-contract=off and correctly rounded
// divide/sqrt; transcendentals come from fp_math.h.  With the operation order
// kept as in the reference, results are reproducible bit for bit.
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
__device__ __forceinline__ bool finite_f(float v) { return (v == v) && (fabsf(v) <= F32_MAX); }
__device__ __forceinline__ float pos_or_zero(float v) { return (finite_f(v) && v > 0.0f) ? v : 0.0f; }
// a*(1-t) + b*t with the product-sum fused, as the reference CPU twin does
// (a.mul_add(1.0 - t, b * t), physics/interpolation.rs:76-79); u = 1 - t.
__device__ __forceinline__ float mix(float a, float b, float t, float u) { return __builtin_fmaf(a, u, b * t); }

// ---------------------------------------------------------------- Philox

struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox10(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 multiply per lane and word (v_mad_u64_u32)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = U4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
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
    const float ru = clampf(unit24(a), 1.1754944e-38f, 0.99999994f);
    const float au = clampf(unit24(b), 0.0f, 0.99999994f);
    const float rad = fpm_sqrtf(-2.0f * fpm_logf_unit(ru));  // ru in [2^-126, 1)
    float sn, cs;
    fpm_sincosf(6.2831855f * au, &sn, &cs);
    g0 = rad * cs;
    g1 = rad * sn;
}

// ---------------------------------------------------------------- met sampling

struct Axis { uint32_t lo, hi; float t; };

__device__ __forceinline__ Axis axis_of(float coord, uint32_t n)
{
    const float c = clampf(coord, 0.0f, (float)(n - 1));
    const float fl = floorf(c);
    const uint32_t l = (uint32_t)fl;
    return Axis{l, min(l + 1u, n - 1u), c - fl};
}

// z [m] -> fractional level.  First k >= 1 with z <= h[k] by bisection over
// the non-decreasing level table in LDS (same result as the linear scan of
// advection_dual_wind.wgsl:94-117).
__device__ __forceinline__ float level_of(float z, const float *lh, uint32_t nz, bool heights)
{
    if (!heights) return z;
    const float h0 = lh[0];
    if (z <= h0) return h0 > 0.0f ? z / h0 : 0.0f;
    uint32_t lo = 1, hi = nz;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (z <= lh[mid]) hi = mid; else lo = mid + 1;
    }
    if (lo >= nz) return (float)(nz - 1);
    const float hp = lh[lo - 1];
    const float den = lh[lo] - hp;
    return den > 0.0f ? (float)(lo - 1) + (z - hp) / den : (float)(lo - 1);
}

struct Wind { float u, v, w; };

// sample_dual_u/v/w (advection_dual_wind.wgsl:142-236): trilinear per time
// level, then the temporal blend.  The 8 corners come from packed 32-byte
// records {u0,v0,w0,u1,v1,w1,-,-}: two 16-byte loads per corner.
__device__ __forceinline__ Wind sample_wind(const StepArgs &a, const float *lh, float alpha, float x,
                                            float y, float z)
{
    const Axis ax = axis_of(x, a.nx);
    const Axis ay = axis_of(y, a.ny);
    const Axis az = axis_of(level_of(z, lh, a.nz, a.has_heights != 0), a.nz);
    const uint32_t ny = a.ny, nz = a.nz;
    const uint32_t r00 = (ax.lo * ny + ay.lo) * nz, r10 = (ax.hi * ny + ay.lo) * nz;
    const uint32_t r01 = (ax.lo * ny + ay.hi) * nz, r11 = (ax.hi * ny + ay.hi) * nz;
    const uint32_t idx[8] = {r00 + az.lo, r10 + az.lo, r01 + az.lo, r11 + az.lo,
                             r00 + az.hi, r10 + az.hi, r01 + az.hi, r11 + az.hi};
    float c[8][6];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float4 p = a.met[2u * idx[k]];
        const float4 q = a.met[2u * idx[k] + 1u];
        c[k][0] = p.x; c[k][1] = p.y; c[k][2] = p.z;
        c[k][3] = p.w; c[k][4] = q.x; c[k][5] = q.y;
    }
    const float ux = 1.0f - ax.t, uy = 1.0f - ay.t, uz = 1.0f - az.t, ua = 1.0f - alpha;
    float v[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const float x0 = mix(c[0][f], c[1][f], ax.t, ux), x1 = mix(c[2][f], c[3][f], ax.t, ux);
        const float x2 = mix(c[4][f], c[5][f], ax.t, ux), x3 = mix(c[6][f], c[7][f], ax.t, ux);
        v[f] = mix(mix(x0, x1, ay.t, uy), mix(x2, x3, ay.t, uy), az.t, uz);
    }
    return Wind{mix(v[0], v[3], alpha, ua), mix(v[1], v[4], alpha, ua), mix(v[2], v[5], alpha, ua)};
}

__device__ __forceinline__ void split_cell(float c, uint32_t n, int32_t &cell, float &frac)
{
    const float cl = clampf(c, 0.0f, (float)(n - 1));
    uint32_t ci = (uint32_t)floorf(cl);
    float fr = cl - (float)ci;
    if (ci >= n - 1u) { ci = n - 1u; fr = 0.0f; }
    cell = (int32_t)ci;
    frac = fr;
}

// ---------------------------------------------------------------- Hanna

// The turbulence regime of a particle is fixed for the whole step: it depends
// on L and the mixing height of its PBL cell only (langevin_fused.wgsl:236).
struct Turb {
    float ust, wst, ol, h;
    bool neu, uns;  // neutral, unstable (else stable)
    float sigu, sigv, sigw, dsigwdz, tlu, tlv, tlw;
};

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
struct VRaw { float sigw, sigw_pre_floor, dsigwdz, tlw, d0, d2; };

__device__ __forceinline__ VRaw vertical_raw(const Turb &t, float z, float zeta, double log_1em3)
{
    const float ust = t.ust, wst = t.wst, ol = t.ol, h = t.h;
    const bool neu = t.neu, uns = t.uns;
    VRaw v;
    // d0: neutral z/u* (corr), unstable z/L, stable -1.3 u* / h
    v.d0 = (uns || neu ? z : -1.3f * ust) / (neu ? ust : (uns ? ol : h));
    // E: neutral e^{-2e-4 corr}, unstable e^{-5 zeta}
    const float E = fpm_expf(neu ? -2.0e-4f * v.d0 : -5.0f * zeta);
    float pa = 0.0f, pb = 0.0f;
    if (!neu) {
        const double L = fpm_log_d((double)(zeta > 0.0f ? zeta : 1.0f));
        pa = zeta > 0.0f ? fpm_exp_d_to_f((double)(uns ? 2.0f / 3.0f : 0.8f) * L) : 0.0f;
        if (uns) pb = fpm_exp_d_to_f((double)(-1.0f / 3.0f) * (zeta >= 1.0e-3f ? L : log_1em3));
    }
    float sigw, sraw = 0.0f;
    if (neu) {
        sraw = 1.3f * ust * E;
        sigw = sraw + SIGMA_FLOOR;
    } else if (uns) {
        const float sq = 1.2f * wst * wst * (1.0f - 0.9f * zeta) * pa + (1.8f - 1.4f * zeta) * ust * ust;
        sigw = fpm_sqrtf(sq) + SIGMA_FLOOR;
    } else {
        sigw = SIGMA_FLOOR + 1.3f * ust * (1.0f - zeta);
    }
    v.sigw_pre_floor = sraw;
    // d1: neutral 0.5 z / sigma_w, unstable 0.5 / sigma_w, stable 0.1 h / sigma_w
    const float d1 = (neu ? 0.5f * z : (uns ? 0.5f : 0.1f * h)) / sigw;
    // d2: neutral d1 / (1 + 1.5e-3 corr), unstable d1 / h
    v.d2 = d1 / (neu ? 1.0f + 1.5e-3f * v.d0 : h);
    if (neu) {
        v.dsigwdz = -2.0e-4f * sraw;
        v.tlw = v.d2;
    } else if (uns) {
        v.dsigwdz = v.d2 * (-1.4f * ust * ust + wst * wst * (0.8f * pb - 1.8f * pa));
        const bool ca = z < fabsf(ol);
        const bool cb = !ca && zeta < 0.1f;
        const float d3 = (ca ? 0.1f * z : (cb ? 0.59f * z : 0.15f * h)) /
                         (ca ? sigw * (0.55f - 0.38f * fabsf(v.d0)) : sigw);
        v.tlw = (ca || cb) ? d3 : d3 * (1.0f - E);
    } else {
        v.dsigwdz = v.d0;
        v.tlw = d1 * pa;
    }
    v.sigw = sigw;
    return v;
}

// hanna_short_w: the floors of langevin_fused.wgsl:335-339
__device__ __forceinline__ void vertical_profile(const Turb &t, float z_in, double log_1em3, float &sigw,
                                                 float &dsigwdz, float &tlw)
{
    const float z = clampf(z_in, 0.0f, t.h);
    const float zeta = clampf(z / t.h, 0.0f, 1.0f);
    const VRaw v = vertical_raw(t, z, zeta, log_1em3);
    tlw = fmaxf(v.tlw, TLW_MIN);
    dsigwdz = v.dsigwdz == 0.0f ? 1.0e-10f : v.dsigwdz;
    sigw = (!finite_f(v.sigw) || v.sigw <= 0.0f) ? SIGMA_FLOOR : v.sigw;
    if (!finite_f(tlw)) tlw = TLW_MIN;
    if (!finite_f(dsigwdz)) dsigwdz = 1.0e-10f;
}

// compute_hanna (langevin_fused.wgsl:224-293)
__device__ __forceinline__ Turb hanna_full(float ust_in, float wst_in, float ol, float h_in, float z_in,
                                           double log_1em3)
{
    Turb t;
    t.ust = fmaxf(ust_in, USTAR_MIN);
    t.wst = fmaxf(wst_in, 0.0f);
    t.h = fmaxf(h_in, H_MIN);
    t.ol = ol;
    t.neu = neutral_regime(ol, t.h);
    t.uns = !t.neu && ol < 0.0f;
    const float ust = t.ust, h = t.h;
    const float z = clampf(z_in, 0.0f, h);
    const float zeta = clampf(z / h, 0.0f, 1.0f);
    const VRaw v = vertical_raw(t, z, zeta, log_1em3);
    float sigu, sigv, tlu, tlv;
    if (t.neu) {
        sigu = SIGMA_FLOOR + 2.0f * ust * fpm_expf(-3.0e-4f * v.d0);
        sigv = v.sigw;
    } else if (t.uns) {
        sigu = SIGMA_FLOOR + ust * fpm_powf(12.0f - 0.5f * h / ol, 1.0f / 3.0f);
        sigv = sigu;
    } else {
        sigu = SIGMA_FLOOR + 2.0f * ust * (1.0f - zeta);
        sigv = SIGMA_FLOOR + 1.3f * ust * (1.0f - zeta);
    }
    // unstable and stable: 0.15 h / sigma_u; neutral: T_Lu = T_Lw
    const float d5 = (0.15f * h) / sigu;
    if (t.neu) {
        tlu = v.d2;
        tlv = tlu;
    } else if (t.uns) {
        tlu = d5;
        tlv = tlu;
    } else {
        tlu = d5 * fpm_sqrtf(zeta);
        tlv = 0.467f * tlu;
    }
    float sigw = v.sigw, dsigwdz = v.dsigwdz, tlw = v.tlw;
    tlu = fmaxf(tlu, TLU_MIN);
    tlv = fmaxf(tlv, TLV_MIN);
    tlw = fmaxf(tlw, TLW_MIN);
    if (dsigwdz == 0.0f) dsigwdz = 1.0e-10f;
    if (!finite_f(sigu) || sigu <= 0.0f) sigu = SIGMA_FLOOR;
    if (!finite_f(sigv) || sigv <= 0.0f) sigv = SIGMA_FLOOR;
    if (!finite_f(sigw) || sigw <= 0.0f) sigw = SIGMA_FLOOR;
    if (!finite_f(tlu)) tlu = TLU_MIN;
    if (!finite_f(tlv)) tlv = TLV_MIN;
    if (!finite_f(tlw)) tlw = TLW_MIN;
    if (!finite_f(dsigwdz)) dsigwdz = 1.0e-10f;
    t.sigu = sigu; t.sigv = sigv; t.sigw = sigw; t.dsigwdz = dsigwdz;
    t.tlu = tlu; t.tlv = tlv; t.tlw = tlw;
    return t;
}

// ---------------------------------------------------------------- Langevin
//
// update_horizontal_component / update_vertical_component
// (langevin_fused.wgsl:350-389).  Both branches of the reference,
//   r <  0.5: (1-r) u + eta sigma sqrt(2r)          [+ dt drift]
//   r >= 0.5: e^{-r} u + eta sigma sqrt(1-e^{-2r})  [+ T_L (1-e^{-r}) drift]
// are one expression with selected coefficients: one division, one exp and
// one sqrt per update whichever branch a lane takes.

__device__ __forceinline__ float ou_horizontal(float prev, float sigma_in, float tl_in, float dt, float eta)
{
    const float sigma = pos_or_zero(sigma_in);
    const float tl = pos_or_zero(tl_in);
    const float r = dt / tl;
    const bool small = r < 0.5f;
    const float e = fpm_expf(-r);
    const float root = fpm_sqrtf(small ? fmaxf(2.0f * r, 0.0f) : fmaxf(1.0f - e * e, 0.0f));
    const float nx = (small ? 1.0f - r : e) * prev + eta * sigma * root;
    return (tl == 0.0f || !finite_f(nx)) ? 0.0f : nx;
}

__device__ __forceinline__ float ou_vertical(float prev, float sigw_in, float dsigwdz, float tlw_in,
                                             float dt, float rho, float eta)
{
    const float sw = pos_or_zero(sigw_in);
    const float tl = pos_or_zero(tlw_in);
    const float drift = sw * (dsigwdz + rho * sw);
    const float r = dt / tl;
    const bool small = r < 0.5f;
    const float e = fpm_expf(-r);
    const float root = fpm_sqrtf(small ? fmaxf(2.0f * r, 0.0f) : fmaxf(1.0f - e * e, 0.0f));
    const float nx = (small ? 1.0f - r : e) * prev + eta * sw * root +
                     (small ? dt * drift : tl * (1.0f - e) * drift);
    return (tl == 0.0f || !finite_f(nx)) ? 0.0f : nx;
}

// PBL reflection (langevin_fused.wgsl:481-500)
__device__ __forceinline__ void reflect(float &pz, float &tw, int32_t &cbt, float zmin, float h)
{
    if (!(h > zmin)) return;
    float z = pz;
    for (int bounce = 0; bounce < 10; ++bounce) {
        if (z < zmin) { z = 2.0f * zmin - z; cbt = -cbt; tw = -tw; }
        else if (z > h) { z = 2.0f * h - z; cbt = -cbt; tw = -tw; }
        else break;
    }
    pz = clampf(z, zmin, h);
}

// ---------------------------------------------------------------- gridding

__device__ __forceinline__ uint32_t grid_cell(const GridArgs &g, const float *oh, int32_t cx, int32_t cy,
                                              float px, float py, float z)
{
    int32_t fx = (int32_t)floorf(px * (float)g.refine_x);
    int32_t fy = (int32_t)floorf(py * (float)g.refine_y);
    fx = max(0, min(fx, (int32_t)g.refine_x - 1));
    fy = max(0, min(fy, (int32_t)g.refine_y - 1));
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

// Wave-level pre-reduction of same-cell contributions followed by one pair of
// global atomics per distinct cell in the wave.  Every lane of the wave must
// call this (lanes without a contribution pass valid = false).  Integer sums
// wrap modulo 2^32 exactly as the reference's atomicAdd does, so the result
// does not depend on the order of reduction.
__device__ __forceinline__ void grid_reduce_wave(const GridArgs &g, bool valid, uint32_t cell, uint32_t q)
{
    uint64_t pending = __ballot(valid);
    const uint32_t lane = __lane_id();
    while (pending) {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)pending) - 1u;
        const uint32_t lc = __shfl(cell, (int)leader);
        const bool mine = valid && cell == lc;
        const uint64_t same = __ballot(mine);
        uint32_t v = mine ? q : 0u;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        if (lane == leader) {
            atomicAdd(&g.counts[lc], (uint32_t)__popcll(same));
            atomicAdd(&g.mass_fp[lc], v);
        }
        pending &= ~same;
    }
}

// NSUB > 0 specialises the vertical sub-step count at compile time (the
// production value is 4, timeloop.rs:1111); NSUB < 0 reads it at run time.
// The kernel is latency-bound at low occupancy; the register budget is
// capped for 3 wavefronts per SIMD (<= 168 VGPRs), the most that fits
// without spilling to scratch.
// XCD-aware chunking: the hardware deals blocks round-robin over the 8
// XCDs; give each XCD a contiguous run of particle chunks so that
// Morton-neighbouring particles share one L2.
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t b, uint32_t nb)
{
    const uint32_t xcd = b % kXcds, k = b / kXcds, per = nb / kXcds, rem = nb % kXcds;
    return xcd < rem ? xcd * (per + 1u) + k : rem * (per + 1u) + (xcd - rem) * per + k;
}

// One block's chunk of the step: slots first + r*BLOCK + threadIdx.x for
// r < a.ppt, temporal weight alpha and Philox base counter ctr.
template <uint32_t BLOCK, int NSUB>
__device__ __forceinline__ void step_chunk(const StepArgs &a, uint64_t first, float alpha, U4 ctr,
                                           const float *lh, const float *oh)
{
    const uint32_t nsub = NSUB >= 0 ? (uint32_t)NSUB : a.n_substeps;
    const Soa s = a.s;
    for (uint32_t r = 0; r < a.ppt; ++r) {
        const uint64_t i = first + (uint64_t)r * BLOCK + threadIdx.x;
        bool live = i < a.n;
        uint32_t flags = 0;
        if (live) flags = s.u(F_FLAGS)[i];
        live = live && (flags & FLAG_ACTIVE);
        bool grid_valid = false;
        uint32_t grid_c = 0, grid_q = 0;
        if (live) {
            int32_t cx = s.i(F_CELL_X)[i], cy = s.i(F_CELL_Y)[i];
            float px = s.f(F_POS_X)[i], py = s.f(F_POS_Y)[i], pz = s.f(F_POS_Z)[i];
            float tu = s.f(F_TURB_U)[i], tv = s.f(F_TURB_V)[i], tw = s.f(F_TURB_W)[i];
            int32_t cbt = s.i(F_CBT)[i];
            float mass[FP_MAX_SPECIES];
#pragma unroll
            for (uint32_t q = 0; q < FP_MAX_SPECIES; ++q)
                if (q < a.n_species) mass[q] = s.f(F_MASS0 + q)[i];

            // ---- advection (advection_dual_wind.wgsl:254-303)
            const float nxm = (float)(a.nx - 1u), nym = (float)(a.ny - 1u);
            const float x0 = (float)cx + px, y0 = (float)cy + py, z0 = pz;
            const Wind s0 = sample_wind(a, lh, alpha, x0, y0, z0);
            const float u0 = s0.u + tu, v0 = s0.v + tv, w0 = s0.w;
            const float qx = clampf(x0 + (u0 * a.x_scale) * a.dt_adv, 0.0f, nxm);
            const float qy = clampf(y0 + (v0 * a.y_scale) * a.dt_adv, 0.0f, nym);
            const float qz = clampf(z0 + (w0 * a.z_scale) * a.dt_adv, 0.0f, a.zmax);
            const Wind s1 = sample_wind(a, lh, alpha, qx, qy, qz);
            const float u1 = s1.u + tu, v1 = s1.v + tv, w1 = s1.w;
            const float cu = 0.5f * (u0 + u1), cv = 0.5f * (v0 + v1), cw = 0.5f * (w0 + w1);
            const float x1 = clampf(x0 + (cu * a.x_scale) * a.dt_adv, 0.0f, nxm);
            const float y1 = clampf(y0 + (cv * a.y_scale) * a.dt_adv, 0.0f, nym);
            pz = clampf(z0 + (cw * a.z_scale) * a.dt_adv, 0.0f, a.zmax);
            split_cell(x1, a.nx, cx, px);
            split_cell(y1, a.ny, cy, py);

            // ---- PBL lookup + Hanna (langevin_fused.wgsl:414-427)
            const float hx = (float)cx + px, hy = (float)cy + py;
            const uint32_t pi = (uint32_t)floorf(clampf(hx, 0.0f, nxm));
            const uint32_t pj = (uint32_t)floorf(clampf(hy, 0.0f, nym));
            const uint32_t pidx = pi * a.ny + pj;
            const float4 pb = a.pbl[pidx];
            const float ol = fabsf(pb.w) < 1.0e-5f ? __builtin_inff() : 1.0f / pb.w;
            const Turb t = hanna_full(pb.x, pb.y, ol, pb.z, pz, a.log_1em3);

            // ---- Philox blocks per slot (langevin_fused.wgsl:430-451)
            const uint64_t cbase = i * a.blocks_per_particle;
            const U4 b0 = philox10(ctr_add(ctr.x, ctr.y, ctr.z, ctr.w, cbase), a.key0, a.key1);
            U4 b1 = U4{0u, 0u, 0u, 0u};
            if (a.blocks_per_particle > 1u)
                b1 = philox10(ctr_add(ctr.x, ctr.y, ctr.z, ctr.w, cbase + 1u), a.key0, a.key1);
            float eu, ev, ew0, ew1, ew2, ew3;
            gauss_pair(b0.x, b0.y, eu, ev);
            gauss_pair(b0.z, b0.w, ew0, ew1);
            gauss_pair(b1.x, b1.y, ew2, ew3);

            // ---- horizontal OU over the full step
            tu = ou_horizontal(tu, t.sigu, t.tlu, a.dt, eu);
            tv = ou_horizontal(tv, t.sigv, t.tlv, a.dt, ev);

            // ---- vertical OU with sub-steps and reflection
            if (nsub == 0u) {
                tw = ou_vertical(tw, t.sigw, t.dsigwdz, t.tlw, a.dt, a.rho_grad, ew0);
            } else {
                const float dts = a.dt / (float)nsub;
                const float zmin = a.min_height;
                float sigw = t.sigw, dsw = t.dsigwdz, tlw = t.tlw;
#pragma unroll
                for (uint32_t sub = 0; sub < (NSUB > 0 ? (uint32_t)NSUB : nsub); ++sub) {
                    const float eta = sub == 0u ? ew0 : (sub == 1u ? ew1 : (sub == 2u ? ew2 : ew3));
                    tw = ou_vertical(tw, sigw, dsw, tlw, dts, a.rho_grad, eta);
                    pz = pz + tw * dts;
                    reflect(pz, tw, cbt, zmin, t.h);
                    // hanna_short_w at the new height (langevin_fused.wgsl:502-513)
                    if (sub + 1u < nsub) vertical_profile(t, pz, a.log_1em3, sigw, dsw, tlw);
                }
            }

            // ---- dry deposition (dry_deposition.wgsl:70-106)
            if (a.dry_mode != FP_DRY_OFF) {
                const float href = fmaxf(a.dry_href, 0.1f);
                if (!(pz >= 2.0f * href)) {
                    const float vd = a.dry_mode == FP_DRY_PER_SLOT ? a.dry_vd_slots[i] : a.dry_vd;
                    const float ex = -fmaxf(vd, 0.0f) * fabsf(a.dt) / (2.0f * href);
                    const float sv = clampf(fpm_expf(ex), 0.0f, 1.0f);
#pragma unroll
                    for (uint32_t q = 0; q < FP_MAX_SPECIES; ++q)
                        if (q < a.n_species) mass[q] = mass[q] * sv;
                }
            }
            // ---- wet deposition (wet_deposition.wgsl:73-111)
            if (a.wet_mode != FP_WET_OFF) {
                float lam, frac;
                if (a.wet_mode == FP_WET_PER_SLOT) {
                    lam = a.wet_lambda_slots[i];
                    frac = a.wet_fraction_slots[i];
                } else if (a.wet_mode == FP_WET_PRECIP_FIELD) {
                    // below_cloud_scavenging_coefficient_gas (physics/wet_scavenging.rs:213-225)
                    const float pr = pos_or_zero(a.precip[pidx]);
                    lam = 0.0f;
                    if (pr > 0.0f) lam = pos_or_zero(pos_or_zero(a.wet_a) * fpm_powf(pr, pos_or_zero(a.wet_b)));
                    frac = a.wet_fraction;
                } else {
                    lam = a.wet_lambda;
                    frac = a.wet_fraction;
                }
                const float l = fmaxf(lam, 0.0f);
                const float dt = fabsf(a.dt);
                const float fr = clampf(frac, 0.0f, 1.0f);
                float prob = 0.0f;
                if (!(l <= 0.0f || dt <= 0.0f || fr <= 0.0f))
                    prob = clampf((1.0f - fpm_expf(-l * dt)) * fr, 0.0f, 1.0f);
                const float sv = 1.0f - prob;
#pragma unroll
                for (uint32_t q = 0; q < FP_MAX_SPECIES; ++q)
                    if (q < a.n_species) mass[q] = mass[q] * sv;
            }
            if (a.loss_mode == FP_LOSS_MIN_MASS && mass[0] < a.loss_min_mass) flags &= ~FLAG_ACTIVE;

            // ---- store
            s.i(F_CELL_X)[i] = cx; s.i(F_CELL_Y)[i] = cy;
            s.f(F_POS_X)[i] = px; s.f(F_POS_Y)[i] = py; s.f(F_POS_Z)[i] = pz;
            s.f(F_TURB_U)[i] = tu; s.f(F_TURB_V)[i] = tv; s.f(F_TURB_W)[i] = tw;
            s.f(F_VEL_U)[i] = cu; s.f(F_VEL_V)[i] = cv; s.f(F_VEL_W)[i] = cw;
            s.i(F_CBT)[i] = cbt;
            s.u(F_FLAGS)[i] = flags;
#pragma unroll
            for (uint32_t q = 0; q < FP_MAX_SPECIES; ++q)
                if (q < a.n_species) s.f(F_MASS0 + q)[i] = mass[q];

            if (a.grid_fused && (flags & FLAG_ACTIVE)) {
                grid_valid = true;
                grid_c = grid_cell(a.grid, oh, cx, cy, px, py, pz);
                grid_q = quantize_mass(mass[a.grid.species], a.grid.mass_scale);
            }
        }
        if (a.grid_fused) grid_reduce_wave(a.grid, grid_valid, grid_c, grid_q);
    }
}

template <uint32_t BLOCK>
__device__ __forceinline__ void load_tables(const StepArgs &a, float *lh, float *oh)
{
    for (uint32_t k = threadIdx.x; k < a.nz; k += BLOCK) lh[k] = a.level_heights[k];
    if (threadIdx.x < FP_MAX_OUTPUT_LEVELS) oh[threadIdx.x] = a.grid.outheights[threadIdx.x];
    __syncthreads();
}

template <uint32_t BLOCK, int NSUB>
__global__ void __launch_bounds__(BLOCK, 3) step_kernel(StepArgs a)
{
    __shared__ float lh[kMaxLevels];
    __shared__ float oh[FP_MAX_OUTPUT_LEVELS];
    load_tables<BLOCK>(a, lh, oh);
    const uint64_t first = (uint64_t)xcd_chunk(blockIdx.x, gridDim.x) * BLOCK * a.ppt;
    step_chunk<BLOCK, NSUB>(a, first, a.alpha, U4{a.c0, a.c1, a.c2, a.c3}, lh, oh);
}

// ---------------------------------------------------------------- ensembles
//
// All members of an ensemble advance in one launch: a block table maps each
// block to (member, first slot); every member has its own particle range,
// met bracket, Philox key/counter and output grid (EnsMember, device
// memory), and they share the step clock, so one launch (and one captured
// graph node) covers members of any size.

__global__ void ens_clock_kernel(EnsClock *c)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t k = c->next_step;
    c->step = k;
    // interpolation_alpha (timeloop.rs:2086-2120): f64, then f32, clamped
    const double t = c->t0_in_bracket + (double)k * c->dt_seconds;
    const float alpha = (float)(t / c->bracket_seconds);
    c->alpha = fminf(fmaxf(alpha, 0.0f), 1.0f);
    c->next_step = k + 1u;
}

template <uint32_t BLOCK, int NSUB>
__global__ void __launch_bounds__(BLOCK, 3) ens_step_kernel(const EnsMember *members,
                                                            const EnsBlock *blocks, const EnsClock *clock)
{
    __shared__ float lh[kMaxLevels];
    __shared__ float oh[FP_MAX_OUTPUT_LEVELS];
    load_tables<BLOCK>(members[0].a, lh, oh);  // one vertical grid for all members
    const EnsBlock eb = blocks[xcd_chunk(blockIdx.x, gridDim.x)];
    const EnsMember &m = members[eb.member];
    const U4 ctr = ctr_add(m.base_ctr[0], m.base_ctr[1], m.base_ctr[2], m.base_ctr[3],
                           (uint64_t)clock->step * m.ctr_per_step);
    step_chunk<BLOCK, NSUB>(m.a, eb.first, clock->alpha, ctr, lh, oh);
}

template <uint32_t BLOCK>
__global__ void __launch_bounds__(BLOCK) grid_kernel(Soa s, uint32_t n, GridArgs g)
{
    __shared__ float oh[FP_MAX_OUTPUT_LEVELS];
    if (threadIdx.x < FP_MAX_OUTPUT_LEVELS) oh[threadIdx.x] = g.outheights[threadIdx.x];
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    bool valid = false;
    uint32_t cell = 0, q = 0;
    if (i < n && (s.u(F_FLAGS)[i] & FLAG_ACTIVE)) {
        valid = true;
        cell = grid_cell(g, oh, s.i(F_CELL_X)[i], s.i(F_CELL_Y)[i], s.f(F_POS_X)[i], s.f(F_POS_Y)[i],
                         s.f(F_POS_Z)[i]);
        q = quantize_mass(s.f(F_MASS0 + g.species)[i], g.mass_scale);
    }
    grid_reduce_wave(g, valid, cell, q);
}

template <uint32_t BLOCK>
hipError_t launch_block(const StepArgs &a, uint32_t grid, hipStream_t st)
{
    if (a.n_substeps == 4u) step_kernel<BLOCK, 4><<<grid, BLOCK, 0, st>>>(a);
    else step_kernel<BLOCK, -1><<<grid, BLOCK, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_step(const StepArgs &a, uint32_t block, hipStream_t st)
{
    if (a.n == 0) return hipSuccess;
    const uint64_t per_block = (uint64_t)block * a.ppt;
    const uint32_t grid = (uint32_t)((a.n + per_block - 1) / per_block);
    switch (block) {
    case 64: return launch_block<64>(a, grid, st);
    case 128: return launch_block<128>(a, grid, st);
    case 256: return launch_block<256>(a, grid, st);
    case 512: return launch_block<512>(a, grid, st);
    case 1024: return launch_block<1024>(a, grid, st);
    default: return hipErrorInvalidValue;
    }
}

template <uint32_t BLOCK>
hipError_t launch_ens_block(const EnsMember *members, const EnsBlock *blocks, EnsClock *clock,
                            uint32_t nblocks, bool sub4, hipStream_t st)
{
    ens_clock_kernel<<<1, 64, 0, st>>>(clock);
    if (nblocks == 0) return hipGetLastError();
    if (sub4) ens_step_kernel<BLOCK, 4><<<nblocks, BLOCK, 0, st>>>(members, blocks, clock);
    else ens_step_kernel<BLOCK, -1><<<nblocks, BLOCK, 0, st>>>(members, blocks, clock);
    return hipGetLastError();
}

hipError_t launch_ens_step(const EnsMember *members, const EnsBlock *blocks, EnsClock *clock,
                           uint32_t nblocks, uint32_t block, bool sub4, hipStream_t st)
{
    switch (block) {
    case 64: return launch_ens_block<64>(members, blocks, clock, nblocks, sub4, st);
    case 128: return launch_ens_block<128>(members, blocks, clock, nblocks, sub4, st);
    case 256: return launch_ens_block<256>(members, blocks, clock, nblocks, sub4, st);
    case 512: return launch_ens_block<512>(members, blocks, clock, nblocks, sub4, st);
    case 1024: return launch_ens_block<1024>(members, blocks, clock, nblocks, sub4, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_grid(const Soa &s, uint32_t n, const GridArgs &g, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const uint32_t grid = (n + 255u) / 256u;
    grid_kernel<256><<<grid, 256, 0, st>>>(s, n, g);
    return hipGetLastError();
}

}  // namespace fp
