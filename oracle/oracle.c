/*
 * oracle.c - CPU restatement of FLEXPART-GPU's production per-particle
 * timestep.  TEST INFRASTRUCTURE ONLY (the checker of the HIP path; see
 * oracle.h).  Scalar f32, one particle at a time, in the operation order of
 * the reference shaders so that every float result is reproducible.
 *
 * Math: built twice by oracle/Makefile.
 *   liboracle.so       uses fp_math.h (deterministic; the HIP kernel uses the
 *                      same primitives, so HIP-vs-oracle parity is bitwise)
 *   liboracle_libm.so  uses glibc expf/logf/sinf/cosf/powf, the functions the
 *                      reference's Rust CPU path resolves to; used to pin the
 *                      deterministic build to libm within tolerance.
 * Both are compiled with -ffp-contract=off.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fp_math.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_LIBM
#define O_EXP(x) expf(x)
#define O_LOG(x) logf(x)
#define O_POW(x, y) powf((x), (y))
static inline void o_sincos(float x, float *s, float *c) { *s = sinf(x); *c = cosf(x); }
#else
#define O_EXP(x) fpm_expf(x)
#define O_LOG(x) fpm_logf(x)
#define O_POW(x, y) fpm_powf((x), (y))
static inline void o_sincos(float x, float *s, float *c) { fpm_sincosf(x, s, c); }
#endif

static inline float o_sqrt(float x) { return sqrtf(x); }
static inline float o_min(float a, float b) { return fminf(a, b); }
static inline float o_max(float a, float b) { return fmaxf(a, b); }
static inline float o_clamp(float x, float lo, float hi) { return o_min(o_max(x, lo), hi); }

#define FLAG_ACTIVE 1u

/* ---------------------------------------------------------------------- */
/* Philox4x32-10: physics/rng.rs:14-87                                      */
/* ---------------------------------------------------------------------- */

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = counter[0], c1 = counter[1], c2 = counter[2], c3 = counter[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)PHILOX_M0 * c0; /* rng.rs:29 */
        const uint64_t p1 = (uint64_t)PHILOX_M1 * c2; /* rng.rs:30 */
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0; /* raise_key rng.rs:40-45 */
        k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 128-bit wrapping add of a 64-bit block offset: rng.rs:69-87. */
void orc_counter_add(const uint32_t counter[4], uint64_t offset, uint32_t out[4])
{
    const uint32_t low = (uint32_t)offset, high = (uint32_t)(offset >> 32);
    uint32_t c0 = counter[0] + low;
    uint32_t carry0 = c0 < counter[0];
    uint32_t c1a = counter[1] + high;
    uint32_t carry1a = c1a < counter[1];
    uint32_t c1 = c1a + carry0;
    uint32_t carry1b = c1 < c1a;
    uint32_t carry1 = carry1a | carry1b;
    uint32_t c2 = counter[2] + carry1;
    uint32_t carry2 = c2 < counter[2];
    uint32_t c3 = counter[3] + carry2;
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_philox_stream(const uint32_t counter[4], const uint32_t key[2], uint64_t n_blocks,
                       uint32_t *out)
{
    for (uint64_t i = 0; i < n_blocks; ++i) {
        uint32_t c[4];
        orc_counter_add(counter, i, c);
        orc_philox4x32_10(c, key, out + 4 * i);
    }
}

/* rng.rs:94-96 */
float orc_u32_to_uniform01(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }

/* langevin_fused.wgsl:186-192 */
void orc_box_muller(float u0, float u1, float out[2])
{
    const float ru = o_clamp(u0, 1.1754944e-38f, 0.99999994f);
    const float au = o_clamp(u1, 0.0f, 0.99999994f);
    const float radius = o_sqrt(-2.0f * O_LOG(ru));
    const float theta = 6.2831855f * au;
    float s, c;
    o_sincos(theta, &s, &c);
    out[0] = radius * c;
    out[1] = radius * s;
}

/* ---------------------------------------------------------------------- */
/* Advection: advection_dual_wind.wgsl:82-303                               */
/* ---------------------------------------------------------------------- */

typedef struct {
    const orc_met *m;
    float alpha;
    int has_heights;
} adv_ctx;

/* advection_dual_wind.wgsl:94-117, generalised from 16 to nz levels */
static float meters_to_grid_level(const adv_ctx *a, float z_m)
{
    const orc_met *m = a->m;
    if (!a->has_heights) return z_m;
    const float h0 = m->level_heights[0];
    if (z_m <= h0) {
        if (h0 > 0.0f) return z_m / h0;
        return 0.0f;
    }
    for (uint32_t k = 1; k < m->nz; ++k) {
        const float hk = m->level_heights[k];
        if (z_m <= hk) {
            const float hprev = m->level_heights[k - 1];
            const float denom = hk - hprev;
            if (denom > 0.0f) return (float)(k - 1) + (z_m - hprev) / denom;
            return (float)(k - 1);
        }
    }
    return (float)(m->nz - 1);
}

static float max_domain_height(const adv_ctx *a)
{
    if (!a->has_heights) return (float)(a->m->nz - 1);
    return a->m->level_heights[a->m->nz - 1];
}

/* a*(1-t) + b*t with the product-sum fused as the reference's CPU twin does:
 * a.mul_add(1.0 - t, b * t) (physics/interpolation.rs:76-79). */
static inline float lerp(float a, float b, float t) { return __builtin_fmaf(a, 1.0f - t, b * t); }

/* clamp_axis advection_dual_wind.wgsl:130-138 */
static void clamp_axis(float coord, uint32_t n, uint32_t *lo, uint32_t *hi, float *frac)
{
    const float maxi = (float)(n - 1);
    const float c = o_clamp(coord, 0.0f, maxi);
    const float lf = floorf(c);
    const uint32_t l = (uint32_t)lf;
    *lo = l;
    *hi = (l + 1 < n - 1) ? l + 1 : n - 1;
    *frac = c - lf;
}

static float trilinear(const float *f, const orc_met *m, const uint32_t ix[2],
                       const uint32_t iy[2], const uint32_t iz[2], float tx, float ty, float tz)
{
#define IDX(a, b, c) ((((uint64_t)ix[a] * m->ny) + iy[b]) * m->nz + iz[c])
    const float c000 = f[IDX(0, 0, 0)], c100 = f[IDX(1, 0, 0)];
    const float c010 = f[IDX(0, 1, 0)], c110 = f[IDX(1, 1, 0)];
    const float c001 = f[IDX(0, 0, 1)], c101 = f[IDX(1, 0, 1)];
    const float c011 = f[IDX(0, 1, 1)], c111 = f[IDX(1, 1, 1)];
#undef IDX
    return lerp(lerp(lerp(c000, c100, tx), lerp(c010, c110, tx), ty),
                lerp(lerp(c001, c101, tx), lerp(c011, c111, tx), ty), tz);
}

/* sample_dual_u/v/w advection_dual_wind.wgsl:142-236 */
static void sample_dual(const adv_ctx *a, float x, float y, float z, float out[3])
{
    const orc_met *m = a->m;
    uint32_t ix[2], iy[2], iz[2];
    float tx, ty, tz;
    clamp_axis(x, m->nx, &ix[0], &ix[1], &tx);
    clamp_axis(y, m->ny, &iy[0], &iy[1], &ty);
    clamp_axis(meters_to_grid_level(a, z), m->nz, &iz[0], &iz[1], &tz);
    const float *f0[3] = {m->u0, m->v0, m->w0};
    const float *f1[3] = {m->u1, m->v1, m->w1};
    for (int c = 0; c < 3; ++c) {
        const float v0 = trilinear(f0[c], m, ix, iy, iz, tx, ty, tz);
        const float v1 = trilinear(f1[c], m, ix, iy, iz, tx, ty, tz);
        out[c] = lerp(v0, v1, a->alpha);
    }
}

/* split_horizontal_coord advection_dual_wind.wgsl:238-248 */
static void split_horizontal(float coord, uint32_t n, int32_t *cell, float *frac)
{
    const float c = o_clamp(coord, 0.0f, (float)(n - 1));
    uint32_t ci = (uint32_t)floorf(c);
    float fr = c - (float)ci;
    if (ci >= n - 1) {
        ci = n - 1;
        fr = 0.0f;
    }
    *cell = (int32_t)ci;
    *frac = fr;
}

/* advection_dual_wind.wgsl:254-303 (main) */
static void advect(orc_particle *p, const adv_ctx *a, const orc_step_params *sp)
{
    const orc_met *m = a->m;
    const float dt = sp->direction < 0 ? -sp->dt_seconds : sp->dt_seconds;
    const float nxm = (float)(m->nx - 1), nym = (float)(m->ny - 1);
    const float zmax = max_domain_height(a);
    const float x0 = (float)p->cell_x + p->pos_x;
    const float y0 = (float)p->cell_y + p->pos_y;
    const float z0 = p->pos_z;
    float s0[3], s1[3];
    sample_dual(a, x0, y0, z0, s0);
    const float u0 = s0[0] + p->turb_u, v0 = s0[1] + p->turb_v, w0 = s0[2];
    const float px = o_clamp(x0 + (u0 * sp->x_scale) * dt, 0.0f, nxm);
    const float py = o_clamp(y0 + (v0 * sp->y_scale) * dt, 0.0f, nym);
    const float pz = o_clamp(z0 + (w0 * sp->z_scale) * dt, 0.0f, zmax);
    sample_dual(a, px, py, pz, s1);
    const float u1 = s1[0] + p->turb_u, v1 = s1[1] + p->turb_v, w1 = s1[2];
    const float cu = 0.5f * (u0 + u1), cv = 0.5f * (v0 + v1), cw = 0.5f * (w0 + w1);
    const float x1 = o_clamp(x0 + (cu * sp->x_scale) * dt, 0.0f, nxm);
    const float y1 = o_clamp(y0 + (cv * sp->y_scale) * dt, 0.0f, nym);
    const float z1 = o_clamp(z0 + (cw * sp->z_scale) * dt, 0.0f, zmax);
    split_horizontal(x1, m->nx, &p->cell_x, &p->pos_x);
    split_horizontal(y1, m->ny, &p->cell_y, &p->pos_y);
    p->pos_z = z1;
    p->vel_u = cu;
    p->vel_v = cv;
    p->vel_w = cw;
}

/* ---------------------------------------------------------------------- */
/* Hanna: langevin_fused.wgsl:224-344 (CPU twin physics/hanna.rs:49-163)    */
/* ---------------------------------------------------------------------- */

#define USTAR_MIN 1.0e-4f
#define TLU_MIN 10.0f
#define TLV_MIN 10.0f
#define TLW_MIN 30.0f
#define SIGMA_FLOOR 1.0e-2f
#define H_MIN 1.0f

static inline int is_finite_f(float v) { return (v == v) && (fabsf(v) <= FLT_MAX); }

typedef struct {
    float ust, wst, ol, h, zeta, sigu, sigv, sigw, dsigwdz, dsigw2dz, tlu, tlv, tlw;
} hanna_t;

static int is_neutral(float ol, float h)
{
    return !is_finite_f(ol) || ol == 0.0f || h / fabsf(ol) < 1.0f;
}

static hanna_t compute_hanna(float ust_in, float wst_in, float ol, float h_in, float z_in)
{
    hanna_t r;
    const float ust = o_max(ust_in, USTAR_MIN);
    const float wst = o_max(wst_in, 0.0f);
    const float h = o_max(h_in, H_MIN);
    const float z = o_clamp(z_in, 0.0f, h);
    const float zeta = o_clamp(z / h, 0.0f, 1.0f);
    float sigu, sigv, sigw, dsigwdz, tlu, tlv, tlw;
    if (is_neutral(ol, h)) {
        const float corr = z / ust;
        sigu = SIGMA_FLOOR + 2.0f * ust * O_EXP(-3.0e-4f * corr);
        sigw = 1.3f * ust * O_EXP(-2.0e-4f * corr);
        dsigwdz = -2.0e-4f * sigw;
        sigw = sigw + SIGMA_FLOOR;
        sigv = sigw;
        tlu = 0.5f * z / sigw / (1.0f + 1.5e-3f * corr);
        tlv = tlu;
        tlw = tlu;
    } else if (ol < 0.0f) {
        sigu = SIGMA_FLOOR + ust * O_POW(12.0f - 0.5f * h / ol, 1.0f / 3.0f);
        sigv = sigu;
        const float sigw_sq = 1.2f * wst * wst * (1.0f - 0.9f * zeta) * O_POW(zeta, 2.0f / 3.0f) +
                              (1.8f - 1.4f * zeta) * ust * ust;
        sigw = o_sqrt(sigw_sq) + SIGMA_FLOOR;
        dsigwdz = 0.5f / sigw / h *
                  (-1.4f * ust * ust +
                   wst * wst *
                       (0.8f * O_POW(o_max(zeta, 1.0e-3f), -1.0f / 3.0f) -
                        1.8f * O_POW(zeta, 2.0f / 3.0f)));
        tlu = 0.15f * h / sigu;
        tlv = tlu;
        if (z < fabsf(ol)) {
            tlw = 0.1f * z / (sigw * (0.55f - 0.38f * fabsf(z / ol)));
        } else if (zeta < 0.1f) {
            tlw = 0.59f * z / sigw;
        } else {
            tlw = 0.15f * h / sigw * (1.0f - O_EXP(-5.0f * zeta));
        }
    } else {
        sigu = SIGMA_FLOOR + 2.0f * ust * (1.0f - zeta);
        sigv = SIGMA_FLOOR + 1.3f * ust * (1.0f - zeta);
        sigw = sigv;
        dsigwdz = -1.3f * ust / h;
        tlu = 0.15f * h / sigu * o_sqrt(zeta);
        tlv = 0.467f * tlu;
        tlw = 0.1f * h / sigw * O_POW(zeta, 0.8f);
    }
    tlu = o_max(tlu, TLU_MIN);
    tlv = o_max(tlv, TLV_MIN);
    tlw = o_max(tlw, TLW_MIN);
    if (dsigwdz == 0.0f) dsigwdz = 1.0e-10f;
    if (!is_finite_f(sigu) || sigu <= 0.0f) sigu = SIGMA_FLOOR;
    if (!is_finite_f(sigv) || sigv <= 0.0f) sigv = SIGMA_FLOOR;
    if (!is_finite_f(sigw) || sigw <= 0.0f) sigw = SIGMA_FLOOR;
    if (!is_finite_f(tlu)) tlu = TLU_MIN;
    if (!is_finite_f(tlv)) tlv = TLV_MIN;
    if (!is_finite_f(tlw)) tlw = TLW_MIN;
    if (!is_finite_f(dsigwdz)) dsigwdz = 1.0e-10f;
    r.ust = ust; r.wst = wst; r.ol = ol; r.h = h; r.zeta = zeta;
    r.sigu = sigu; r.sigv = sigv; r.sigw = sigw;
    r.dsigwdz = dsigwdz; r.dsigw2dz = 2.0f * sigw * dsigwdz;
    r.tlu = tlu; r.tlv = tlv; r.tlw = tlw;
    return r;
}

void orc_hanna(float ust, float wst, float ol, float h, float z, float out[13])
{
    const hanna_t r = compute_hanna(ust, wst, ol, h, z);
    const float v[13] = {r.ust, r.wst, r.ol, r.h, r.zeta, r.sigu, r.sigv,
                         r.sigw, r.dsigwdz, r.dsigw2dz, r.tlu, r.tlv, r.tlw};
    memcpy(out, v, sizeof v);
}

/* hanna_short_w langevin_fused.wgsl:296-344 */
static void hanna_short_w(float ust_in, float wst_in, float ol, float h_in, float z_in,
                          float *o_sigw, float *o_dsigwdz, float *o_tlw)
{
    const float ust = o_max(ust_in, USTAR_MIN);
    const float wst = o_max(wst_in, 0.0f);
    const float h = o_max(h_in, H_MIN);
    const float z = o_clamp(z_in, 0.0f, h);
    const float zeta = o_clamp(z / h, 0.0f, 1.0f);
    float sigw, dsigwdz, tlw;
    if (is_neutral(ol, h)) {
        const float corr = z / ust;
        sigw = 1.3f * ust * O_EXP(-2.0e-4f * corr);
        dsigwdz = -2.0e-4f * sigw;
        sigw = sigw + SIGMA_FLOOR;
        tlw = 0.5f * z / sigw / (1.0f + 1.5e-3f * corr);
    } else if (ol < 0.0f) {
        const float sigw_sq = 1.2f * wst * wst * (1.0f - 0.9f * zeta) * O_POW(zeta, 2.0f / 3.0f) +
                              (1.8f - 1.4f * zeta) * ust * ust;
        sigw = o_sqrt(sigw_sq) + SIGMA_FLOOR;
        dsigwdz = 0.5f / sigw / h *
                  (-1.4f * ust * ust +
                   wst * wst *
                       (0.8f * O_POW(o_max(zeta, 1.0e-3f), -1.0f / 3.0f) -
                        1.8f * O_POW(zeta, 2.0f / 3.0f)));
        if (z < fabsf(ol)) {
            tlw = 0.1f * z / (sigw * (0.55f - 0.38f * fabsf(z / ol)));
        } else if (zeta < 0.1f) {
            tlw = 0.59f * z / sigw;
        } else {
            tlw = 0.15f * h / sigw * (1.0f - O_EXP(-5.0f * zeta));
        }
    } else {
        sigw = SIGMA_FLOOR + 1.3f * ust * (1.0f - zeta);
        dsigwdz = -1.3f * ust / h;
        tlw = 0.1f * h / sigw * O_POW(zeta, 0.8f);
    }
    tlw = o_max(tlw, TLW_MIN);
    if (dsigwdz == 0.0f) dsigwdz = 1.0e-10f;
    if (!is_finite_f(sigw) || sigw <= 0.0f) sigw = SIGMA_FLOOR;
    if (!is_finite_f(tlw)) tlw = TLW_MIN;
    if (!is_finite_f(dsigwdz)) dsigwdz = 1.0e-10f;
    *o_sigw = sigw;
    *o_dsigwdz = dsigwdz;
    *o_tlw = tlw;
}

/* ---------------------------------------------------------------------- */
/* Langevin: langevin_fused.wgsl:198-207, 350-396                            */
/* ---------------------------------------------------------------------- */

static inline float sanitize_positive(float v) { return (is_finite_f(v) && v > 0.0f) ? v : 0.0f; }

static float update_horizontal(float prev, float sigma_in, float tl_in, float dt, float eta)
{
    const float sigma = sanitize_positive(sigma_in);
    const float tl = sanitize_positive(tl_in);
    if (tl == 0.0f) return 0.0f;
    const float ratio = dt / tl;
    float next;
    if (ratio < 0.5f) {
        next = (1.0f - ratio) * prev + eta * sigma * o_sqrt(o_max(2.0f * ratio, 0.0f));
    } else {
        const float corr = O_EXP(-ratio);
        next = corr * prev + eta * sigma * o_sqrt(o_max(1.0f - corr * corr, 0.0f));
    }
    return is_finite_f(next) ? next : 0.0f;
}

static float update_vertical(float prev, const hanna_t *hn, float dt, float rho, float eta)
{
    const float sigma_w = sanitize_positive(hn->sigw);
    const float tlw = sanitize_positive(hn->tlw);
    if (tlw == 0.0f) return 0.0f;
    const float drift = sigma_w * (hn->dsigwdz + rho * sigma_w);
    const float ratio = dt / tlw;
    float next;
    if (ratio < 0.5f) {
        next = (1.0f - ratio) * prev + eta * sigma_w * o_sqrt(o_max(2.0f * ratio, 0.0f)) +
               dt * drift;
    } else {
        const float corr = O_EXP(-ratio);
        next = corr * prev + eta * sigma_w * o_sqrt(o_max(1.0f - corr * corr, 0.0f)) +
               tlw * (1.0f - corr) * drift;
    }
    return is_finite_f(next) ? next : 0.0f;
}

/* PBL reflection, langevin_fused.wgsl:481-500: only when h > z_min; up to 10
 * mirror images about z_min or h, each negating turb_w and cbt, then a clamp
 * to [z_min, h]. */
static void pbl_reflect(float *pz, float *tw, int32_t *pcbt, float z_min, float h)
{
    if (!(h > z_min)) return;
    float z = *pz;
    int32_t cbt = *pcbt;
    for (int bounce = 0; bounce < 10; ++bounce) {
        if (z < z_min) {
            z = 2.0f * z_min - z;
            cbt = -cbt;
            *tw = -*tw;
        } else if (z > h) {
            z = 2.0f * h - z;
            cbt = -cbt;
            *tw = -*tw;
        } else {
            break;
        }
    }
    *pz = o_clamp(z, z_min, h);
    *pcbt = cbt;
}

void orc_reflect(float *z, float *turb_w, int32_t *cbt, float z_min, float h)
{
    pbl_reflect(z, turb_w, cbt, z_min, h);
}

void orc_hanna_short_w(float ust, float wst, float ol, float h, float z, float out[3])
{
    hanna_short_w(ust, wst, ol, h, z, &out[0], &out[1], &out[2]);
}

/* langevin_fused.wgsl:402-518 (main) */
static void langevin(orc_particle *p, uint64_t slot, const orc_pbl *pbl,
                     const orc_step_params *sp)
{
    const float x = (float)p->cell_x + p->pos_x;
    const float y = (float)p->cell_y + p->pos_y;
    const uint32_t i = (uint32_t)floorf(o_clamp(x, 0.0f, (float)(pbl->nx - 1)));
    const uint32_t j = (uint32_t)floorf(o_clamp(y, 0.0f, (float)(pbl->ny - 1)));
    const uint64_t idx = (uint64_t)i * pbl->ny + j;
    const float oli = pbl->oli[idx];
    /* obukhov_length_from_inverse langevin_fused.wgsl:213-218 */
    const float ol = fabsf(oli) < 1.0e-5f ? __builtin_inff() : 1.0f / oli;
    hanna_t hn = compute_hanna(pbl->ustar[idx], pbl->wstar[idx], ol, pbl->hmix[idx], p->pos_z);

    const uint32_t n_vert = sp->n_substeps > 1 ? sp->n_substeps : 1;
    const uint32_t blocks = (2u + n_vert + 3u) / 4u;
    uint32_t c[4], b0[4], b1[4] = {0, 0, 0, 0};
    orc_counter_add(sp->counter, slot * blocks, c);
    orc_philox4x32_10(c, sp->key, b0);
    if (blocks > 1) {
        orc_counter_add(sp->counter, slot * blocks + 1, c);
        orc_philox4x32_10(c, sp->key, b1);
    }
    float euv[2], ew01[2], ew23[2];
    orc_box_muller(orc_u32_to_uniform01(b0[0]), orc_u32_to_uniform01(b0[1]), euv);
    orc_box_muller(orc_u32_to_uniform01(b0[2]), orc_u32_to_uniform01(b0[3]), ew01);
    orc_box_muller(orc_u32_to_uniform01(b1[0]), orc_u32_to_uniform01(b1[1]), ew23);

    const float dt = sp->dt_seconds;
    p->turb_u = update_horizontal(p->turb_u, hn.sigu, hn.tlu, dt, euv[0]);
    p->turb_v = update_horizontal(p->turb_v, hn.sigv, hn.tlv, dt, euv[1]);

    if (sp->n_substeps == 0) {
        p->turb_w = update_vertical(p->turb_w, &hn, dt, sp->rho_grad_over_rho, ew01[0]);
        return;
    }
    const float dt_sub = dt / (float)sp->n_substeps;
    const float h = hn.h;
    const float z_min = sp->min_height_m;
    const float eta_w[4] = {ew01[0], ew01[1], ew23[0], ew23[1]};
    for (uint32_t sub = 0; sub < sp->n_substeps; ++sub) {
        const float eta = eta_w[sub < 3 ? sub : 3];
        p->turb_w = update_vertical(p->turb_w, &hn, dt_sub, sp->rho_grad_over_rho, eta);
        p->pos_z = p->pos_z + p->turb_w * dt_sub;
        pbl_reflect(&p->pos_z, &p->turb_w, &p->cbt, z_min, h);
        if (sub < sp->n_substeps - 1) {
            float sw, dsw, tw;
            hanna_short_w(hn.ust, hn.wst, hn.ol, hn.h, p->pos_z, &sw, &dsw, &tw);
            hn.sigw = sw;
            hn.dsigwdz = dsw;
            hn.dsigw2dz = 2.0f * sw * dsw;
            hn.tlw = tw;
            hn.zeta = o_clamp(p->pos_z / hn.h, 0.0f, 1.0f);
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Deposition: dry_deposition.wgsl:70-106, wet_deposition.wgsl:73-111      */
/* ---------------------------------------------------------------------- */

static void dry_deposition(orc_particle *p, uint64_t slot, const orc_step_params *sp)
{
    const float href = o_max(sp->dry_href, 0.1f);
    if (p->pos_z >= 2.0f * href) return;
    const float vd = sp->dry_mode == 2 ? sp->dry_vd_slots[slot] : sp->dry_vd;
    const float exponent = -o_max(vd, 0.0f) * fabsf(sp->dt_seconds) / (2.0f * href);
    const float survival = o_clamp(O_EXP(exponent), 0.0f, 1.0f);
    for (int s = 0; s < 4; ++s) p->mass[s] = p->mass[s] * survival;
}

/* below_cloud_scavenging_coefficient_gas physics/wet_scavenging.rs:213-225 */
static float scavenging_from_precip(float precip_in, float a_in, float b_in)
{
    const float precip = sanitize_positive(precip_in);
    if (precip <= 0.0f) return 0.0f;
    const float a = sanitize_positive(a_in);
    const float b = sanitize_positive(b_in);
    return sanitize_positive(a * O_POW(precip, b));
}

static void wet_deposition(orc_particle *p, uint64_t slot, const orc_pbl *pbl,
                           const orc_step_params *sp)
{
    float lam, frac;
    if (sp->wet_mode == 2) {
        lam = sp->wet_lambda_slots[slot];
        frac = sp->wet_fraction_slots[slot];
    } else if (sp->wet_mode == 3) {
        const float x = (float)p->cell_x + p->pos_x;
        const float y = (float)p->cell_y + p->pos_y;
        const uint32_t i = (uint32_t)floorf(o_clamp(x, 0.0f, (float)(pbl->nx - 1)));
        const uint32_t j = (uint32_t)floorf(o_clamp(y, 0.0f, (float)(pbl->ny - 1)));
        lam = scavenging_from_precip(pbl->precip[(uint64_t)i * pbl->ny + j], sp->wet_a,
                                     sp->wet_b);
        frac = sp->wet_fraction;
    } else {
        lam = sp->wet_lambda;
        frac = sp->wet_fraction;
    }
    const float lambda = o_max(lam, 0.0f);
    const float dt = fabsf(sp->dt_seconds);
    const float fraction = o_clamp(frac, 0.0f, 1.0f);
    float prob = 0.0f;
    if (!(lambda <= 0.0f || dt <= 0.0f || fraction <= 0.0f)) {
        const float local = 1.0f - O_EXP(-lambda * dt);
        prob = o_clamp(local * fraction, 0.0f, 1.0f);
    }
    const float survival = 1.0f - prob;
    for (int s = 0; s < 4; ++s) p->mass[s] = p->mass[s] * survival;
}

/* Stage entry points for the closed-form unit tests. */
float orc_update_horizontal(float prev, float sigma, float tl, float dt, float eta)
{
    return update_horizontal(prev, sigma, tl, dt, eta);
}

float orc_update_vertical(float prev, float sigw, float dsigwdz, float tlw, float dt, float rho,
                          float eta)
{
    hanna_t hn;
    memset(&hn, 0, sizeof hn);
    hn.sigw = sigw;
    hn.dsigwdz = dsigwdz;
    hn.tlw = tlw;
    return update_vertical(prev, &hn, dt, rho, eta);
}

void orc_advect(orc_particle *p, uint64_t n, const orc_met *met, const orc_step_params *sp)
{
    adv_ctx a;
    a.m = met;
    a.alpha = o_clamp(sp->alpha, 0.0f, 1.0f);
    a.has_heights = met->level_heights && met->level_heights[met->nz - 1] > 0.0f;
    for (uint64_t i = 0; i < n; ++i)
        if (p[i].flags & FLAG_ACTIVE) advect(&p[i], &a, sp);
}

void orc_deposit(orc_particle *p, uint64_t n, const orc_pbl *pbl, const orc_step_params *sp)
{
    for (uint64_t i = 0; i < n; ++i) {
        if ((p[i].flags & FLAG_ACTIVE) == 0u) continue;
        if (sp->dry_mode != 0) dry_deposition(&p[i], i, sp);
        if (sp->wet_mode != 0) wet_deposition(&p[i], i, pbl, sp);
    }
}

/* ---------------------------------------------------------------------- */
/* Step: particle_step.wgsl:684-883 ordering                               */
/* ---------------------------------------------------------------------- */

int orc_step(orc_particle *p, uint64_t n, const orc_met *met, const orc_pbl *pbl,
             const orc_step_params *sp, uint32_t next_counter[4])
{
    adv_ctx a;
    a.m = met;
    a.alpha = o_clamp(sp->alpha, 0.0f, 1.0f); /* gpu/advection.rs:639 */
    a.has_heights = met->level_heights && met->level_heights[met->nz - 1] > 0.0f;
    const int64_t nn = (int64_t)n;
#ifdef _OPENMP
    const int nt = sp->n_threads > 0 ? sp->n_threads : 1;
#pragma omp parallel for schedule(static, 4096) num_threads(nt)
#endif
    for (int64_t i = 0; i < nn; ++i) {
        orc_particle *q = &p[i];
        if ((q->flags & FLAG_ACTIVE) == 0u) continue;
        advect(q, &a, sp);
        langevin(q, (uint64_t)i, pbl, sp);
        if (sp->dry_mode != 0) dry_deposition(q, (uint64_t)i, sp);
        if (sp->wet_mode != 0) wet_deposition(q, (uint64_t)i, pbl, sp);
        if (sp->loss_mode == 1 && q->mass[0] < sp->loss_min_mass) q->flags &= ~FLAG_ACTIVE;
    }
    /* host counter advance gpu/langevin_fused.rs:229-233 */
    const uint64_t per = sp->n_substeps <= 2 ? 1 : 2;
    orc_counter_add(sp->counter, n * per, next_counter);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Gridding: concentration_gridding.wgsl:67-128                             */
/* ---------------------------------------------------------------------- */

void orc_grid_accumulate(const orc_particle *p, uint64_t n, const orc_grid_params *gp,
                         uint32_t *counts, uint32_t *mass_fp)
{
    const int metres = gp->outheights[gp->nz - 1] > 0.0f;
    for (uint64_t i = 0; i < n; ++i) {
        const orc_particle *q = &p[i];
        if ((q->flags & FLAG_ACTIVE) == 0u) continue;
        /* output cell: the met cell (refine 1, the reference mapping) or an
         * integer refinement of it */
        int32_t fx = (int32_t)floorf(q->pos_x * (float)gp->refine_x);
        int32_t fy = (int32_t)floorf(q->pos_y * (float)gp->refine_y);
        if (fx > (int32_t)gp->refine_x - 1) fx = (int32_t)gp->refine_x - 1;
        if (fy > (int32_t)gp->refine_y - 1) fy = (int32_t)gp->refine_y - 1;
        if (fx < 0) fx = 0;
        if (fy < 0) fy = 0;
        int64_t ix = (int64_t)q->cell_x * gp->refine_x + fx;
        int64_t iy = (int64_t)q->cell_y * gp->refine_y + fy;
        if (ix < 0) ix = 0;
        if (ix > (int64_t)gp->nx - 1) ix = gp->nx - 1;
        if (iy < 0) iy = 0;
        if (iy > (int64_t)gp->ny - 1) iy = gp->ny - 1;
        uint32_t iz;
        if (metres) {
            iz = gp->nz - 1;
            for (uint32_t k = 0; k < gp->nz; ++k) {
                if (q->pos_z <= gp->outheights[k]) {
                    iz = k;
                    break;
                }
            }
        } else {
            int32_t l = (int32_t)floorf(q->pos_z);
            if (l < 0) l = 0;
            if (l > (int32_t)gp->nz - 1) l = (int32_t)gp->nz - 1;
            iz = (uint32_t)l;
        }
        const uint64_t cell = ((uint64_t)ix * gp->ny + (uint64_t)iy) * gp->nz + iz;
        counts[cell] += 1u;
        const float m = o_max(q->mass[gp->species], 0.0f);
        uint32_t qv = 0;
        if (m > 0.0f) {
            const float scaled = m * gp->mass_scale + 0.5f;
            qv = scaled >= 4294967295.0f ? 0xffffffffu : (uint32_t)scaled;
        }
        mass_fp[cell] += qv;
    }
}

/* ---------------------------------------------------------------------- */
/* Compaction: compaction.wgsl:65-217                                       */
/* ---------------------------------------------------------------------- */

uint64_t orc_compact(const orc_particle *p, uint64_t n, orc_particle *out,
                     uint32_t *survivor_map)
{
    uint64_t a = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (p[i].flags & FLAG_ACTIVE) {
            if (survivor_map) survivor_map[a] = (uint32_t)i;
            out[a] = p[i];
            ++a;
        }
    }
    for (uint64_t i = a; i < n; ++i) { /* deactivated tail, compaction.wgsl:212-216 */
        out[i] = p[i];
        out[i].flags &= ~FLAG_ACTIVE;
    }
    return a;
}

/* ---------------------------------------------------------------------- */
/* Morton sort: particles/mod.rs:415-441, 616-699                          */
/* ---------------------------------------------------------------------- */

static uint64_t split_by_2_zeros(uint32_t v)
{
    /* spread the low 21 bits so that bit b lands at bit 3b */
    uint64_t x = v & 0x1FFFFFu;
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) r |= ((x >> b) & 1u) << (3 * b);
    return r;
}

uint64_t orc_morton_encode_3d_21bit(uint32_t x, uint32_t y, uint32_t z)
{
    return split_by_2_zeros(x) | (split_by_2_zeros(y) << 1) | (split_by_2_zeros(z) << 2);
}

static uint32_t quantize_axis(double v, double lo, double hi, uint32_t max_bin)
{
    if (!isfinite(v) || !isfinite(lo) || !isfinite(hi)) return 0;
    const double span = hi - lo;
    if (span <= DBL_EPSILON) return 0;
    double t = (v - lo) / span;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return (uint32_t)floor(t * (double)max_bin);
}

static int active_bounds(const orc_particle *p, uint64_t n, double b[6])
{
    int found = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (!(p[i].flags & FLAG_ACTIVE)) continue;
        const double x = (double)p[i].cell_x + (double)p[i].pos_x;
        const double y = (double)p[i].cell_y + (double)p[i].pos_y;
        const double z = (double)p[i].pos_z;
        if (!found) {
            b[0] = b[1] = x; b[2] = b[3] = y; b[4] = b[5] = z;
            found = 1;
        } else {
            b[0] = fmin(b[0], x); b[1] = fmax(b[1], x);
            b[2] = fmin(b[2], y); b[3] = fmax(b[3], y);
            b[4] = fmin(b[4], z); b[5] = fmax(b[5], z);
        }
    }
    return found;
}

void orc_morton_keys(const orc_particle *p, uint64_t n, uint32_t bits, int explicit_bounds,
                     const double bounds[6], uint64_t *keys)
{
    double b[6];
    int have = 1;
    if (explicit_bounds) memcpy(b, bounds, sizeof b);
    else have = active_bounds(p, n, b);
    const uint32_t max_bin = (1u << bits) - 1u;
    for (uint64_t i = 0; i < n; ++i) {
        if (!(p[i].flags & FLAG_ACTIVE) || !have) {
            keys[i] = UINT64_MAX;
            continue;
        }
        const double x = (double)p[i].cell_x + (double)p[i].pos_x;
        const double y = (double)p[i].cell_y + (double)p[i].pos_y;
        const double z = (double)p[i].pos_z;
        keys[i] = orc_morton_encode_3d_21bit(quantize_axis(x, b[0], b[1], max_bin),
                                             quantize_axis(y, b[2], b[3], max_bin),
                                             quantize_axis(z, b[4], b[5], max_bin));
    }
}

/* stable merge sort of indices by key */
static void merge_sort(uint32_t *idx, uint32_t *tmp, const uint64_t *keys, uint64_t n)
{
    for (uint64_t w = 1; w < n; w *= 2) {
        for (uint64_t lo = 0; lo < n; lo += 2 * w) {
            uint64_t mid = lo + w < n ? lo + w : n;
            uint64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            uint64_t a = lo, b = mid, k = lo;
            while (a < mid && b < hi) tmp[k++] = keys[idx[b]] < keys[idx[a]] ? idx[b++] : idx[a++];
            while (a < mid) tmp[k++] = idx[a++];
            while (b < hi) tmp[k++] = idx[b++];
        }
        memcpy(idx, tmp, n * sizeof *idx);
    }
}

int orc_sort_morton(const orc_particle *p, uint64_t n, uint32_t bits, int explicit_bounds,
                    const double bounds[6], orc_particle *out, uint32_t *new_to_old)
{
    if (bits < 1 || bits > 21) return -1; /* validate_bits_per_axis particles/mod.rs:406-413 */
    uint64_t *keys = (uint64_t *)malloc(n * sizeof *keys);
    uint32_t *idx = (uint32_t *)malloc(n * sizeof *idx);
    uint32_t *tmp = (uint32_t *)malloc(n * sizeof *tmp);
    if (!keys || !idx || !tmp) {
        free(keys); free(idx); free(tmp);
        return -2;
    }
    orc_morton_keys(p, n, bits, explicit_bounds, bounds, keys);
    for (uint64_t i = 0; i < n; ++i) idx[i] = (uint32_t)i;
    merge_sort(idx, tmp, keys, n);
    for (uint64_t i = 0; i < n; ++i) {
        out[i] = p[idx[i]];
        if (new_to_old) new_to_old[i] = idx[i];
    }
    free(keys); free(idx); free(tmp);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Math probe for the accuracy tests                                        */
/* ---------------------------------------------------------------------- */

void orc_math_eval(int fn, const float *x, const float *y, float *out, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i) {
        float s, c;
        switch (fn) {
        case 0: out[i] = fpm_expf(x[i]); break;
        case 1: out[i] = expf(x[i]); break;
        case 2: out[i] = fpm_logf(x[i]); break;
        case 3: out[i] = logf(x[i]); break;
        case 4: fpm_sincosf(x[i], &s, &c); out[i] = s; break;
        case 5: out[i] = sinf(x[i]); break;
        case 6: fpm_sincosf(x[i], &s, &c); out[i] = c; break;
        case 7: out[i] = cosf(x[i]); break;
        case 8: out[i] = fpm_powf(x[i], y[i]); break;
        case 9: out[i] = powf(x[i], y[i]); break;
        case 10: out[i] = (float)exp((double)x[i]); break;
        case 11: out[i] = (float)log((double)x[i]); break;
        case 12: out[i] = (float)sin((double)x[i]); break;
        case 13: out[i] = (float)cos((double)x[i]); break;
        case 14: out[i] = (float)pow((double)x[i], (double)y[i]); break;
        default: out[i] = 0.0f;
        }
    }
}

/* meters_to_grid_level over a level table, for the lookup's parity tests */
void orc_level_eval(const float *level_heights, uint32_t nz, const float *z, float *out, uint64_t n)
{
    orc_met m;
    memset(&m, 0, sizeof m);
    m.nz = nz;
    m.level_heights = level_heights;
    adv_ctx a;
    a.m = &m;
    a.alpha = 0.0f;
    a.has_heights = level_heights[nz - 1] > 0.0f;
    for (uint64_t i = 0; i < n; ++i) out[i] = meters_to_grid_level(&a, z[i]);
}

/* ---------------------------------------------------------------- release */

/* target_emitted_count (release/mod.rs:181-197) */
static uint64_t orc_release_target(const orc_release *r, uint64_t released, int64_t now_s)
{
    if (now_s < r->start_s || now_s > r->end_s) return released;
    if (r->start_s == r->end_s || now_s == r->end_s) return r->particle_count;
    const uint64_t elapsed = (uint64_t)(now_s - r->start_s);
    const uint64_t duration = (uint64_t)(r->end_s - r->start_s);
    return (uint64_t)(((unsigned __int128)r->particle_count * elapsed) / duration);
}

/* sample_stratified (release/mod.rs:264-283) */
static void orc_release_sample(const orc_release *r, uint64_t idx, double *lon, double *lat, double *z)
{
    double c = ceil(cbrt((double)r->particle_count));
    if (c < 1.0) c = 1.0;
    const uint64_t nx = (uint64_t)c, ny = nx;
    uint64_t nxy = nx * ny;
    if (nxy < 1) nxy = 1;
    uint64_t nz = (r->particle_count + nxy - 1) / nxy;
    if (nz < 1) nz = 1;
    const uint64_t ix = idx % nx, iy = (idx / nx) % ny, iz = (idx / nxy) % nz;
    const double fx = ((double)ix + 0.5) / (double)nx;
    const double fy = ((double)iy + 0.5) / (double)ny;
    const double fz = ((double)iz + 0.5) / (double)nz;
    *lon = r->lon_min + fx * (r->lon_max - r->lon_min);
    *lat = r->lat_min + fy * (r->lat_max - r->lat_min);
    *z = r->z_min + fz * (r->z_max - r->z_min);
}

uint64_t orc_release_inject(const orc_grid_domain *grid, const orc_release *rel, uint32_t n_rel,
                            uint64_t *released, int64_t now_s, int32_t particle_time,
                            orc_particle *out, uint64_t max_out, int *status)
{
    uint64_t w = 0;
    *status = 0;
    for (uint32_t k = 0; k < n_rel; ++k) {
        const orc_release *r = &rel[k];
        const uint64_t target = orc_release_target(r, released[k], now_s);
        if (target <= released[k]) continue;
        const float mass = (float)(r->mass_kg / (double)r->particle_count);
        for (uint64_t idx = released[k]; idx < target; ++idx) {
            double lon, lat, z;
            orc_release_sample(r, idx, &lon, &lat, &z);
            const double gx = (lon - grid->xlon0) / grid->dx;
            const double gy = (lat - grid->ylat0) / grid->dy;
            if (!(gx >= 0.0 && gx < (double)(grid->nx - 1) && gy >= 0.0 && gy < (double)(grid->ny - 1))) {
                *status = 1;
                return w;
            }
            if (w >= max_out) {
                *status = 2;
                return w;
            }
            orc_particle p;
            memset(&p, 0, sizeof p);
            const double flx = floor(gx), fly = floor(gy);
            p.cell_x = (int32_t)flx;
            p.cell_y = (int32_t)fly;
            p.pos_x = (float)(gx - flx);
            p.pos_y = (float)(gy - fly);
            p.pos_z = (float)z;
            p.flags = FLAG_ACTIVE;
            p.mass[0] = mass;
            p.time = particle_time;
            p.time_mem = particle_time;
            p.release_point = r->release_point;
            p.class_id = r->class_id;
            out[w++] = p;
        }
        released[k] = target;
    }
    return w;
}

/* ---------------------------------------------------------------------- */
/* PBL from surface fields: io/temporal.rs:122-214 + pbl_diagnostics.wgsl   */
/* ---------------------------------------------------------------------- */

#define PD_R_AIR 287.05f /* lib.rs:24 */
#define PD_GA 9.81f      /* lib.rs:26 */
#define PD_CPA 1004.6f   /* lib.rs:28 */
#define PD_KARMAN 0.4f
#define PD_USTAR_MIN 1.0e-4f
#define PD_NEUTRAL_OLI 1.0e-5f

/* is_finite_scalar and the sanitize_* helpers (pbl_diagnostics.wgsl:98-121) */
static int pd_finite(float v) { return v == v && fabsf(v) <= FLT_MAX; }
static float pd_positive(float v, float fb) { return (pd_finite(v) && v > 0.0f) ? v : fb; }
static float pd_nonneg(float v, float fb) { return (pd_finite(v) && v >= 0.0f) ? v : fb; }
static float pd_fin(float v, float fb) { return pd_finite(v) ? v : fb; }

/* estimate_friction_velocity (pbl_diagnostics.wgsl:130-152) */
static float pd_friction_velocity(float stress, float air_density, float u10, float v10, float z0_in,
                                  float zref_in)
{
    const float rho = pd_positive(air_density, 1.225f);
    const float safe_stress = pd_nonneg(stress, 0.0f);
    if (safe_stress > 0.0f) return o_max(o_sqrt(safe_stress / rho), PD_USTAR_MIN);
    const float z0 = pd_positive(z0_in, 0.1f);
    const float zref = o_max(pd_positive(zref_in, 10.0f), 1.01f * z0);
    const float speed = o_max(o_sqrt(u10 * u10 + v10 * v10), 0.0f);
    if (speed <= 0.0f) return PD_USTAR_MIN;
    const float ustar = PD_KARMAN * speed / o_max(O_LOG(zref / z0), 1.0e-6f);
    return o_max(pd_positive(ustar, PD_USTAR_MIN), PD_USTAR_MIN);
}

/* obukhov_length_from_flux (pbl_diagnostics.wgsl:160-187) */
static float pd_obukhov_length(float ustar, float temp_k, float air_density, float heat_flux, float threshold)
{
    const float h = pd_fin(heat_flux, 0.0f);
    if (fabsf(h) < o_max(threshold, 0.0f)) return INFINITY;
    const float u = o_max(pd_positive(ustar, PD_USTAR_MIN), PD_USTAR_MIN);
    const float t = pd_positive(temp_k, 300.0f);
    const float d = pd_positive(air_density, 1.225f);
    const float numerator = -(d * PD_CPA * t * u * u * u);
    const float denominator = PD_KARMAN * PD_GA * h;
    if (fabsf(denominator) < 1.0e-9f) return INFINITY;
    const float ol = numerator / denominator;
    if (!pd_finite(ol) || fabsf(ol) > 1.0e9f) return INFINITY;
    return ol;
}

/* inverse_obukhov_from_length (pbl_diagnostics.wgsl:190-195) */
static float pd_inverse_obukhov(float ol)
{
    if (!pd_finite(ol) || fabsf(ol) > 1.0e5f) return 0.0f;
    return 1.0f / ol;
}

/* clamp_mixing_height (pbl_diagnostics.wgsl:197-199): max(clamp(v, lo, hi), lo) */
static float pd_clamp_hmix(float v, const orc_pbl_options *o)
{
    return o_max(o_min(o_max(v, o->hmix_min_m), o->hmix_max_m), o->hmix_min_m);
}

/* convective_velocity_scale (pbl_diagnostics.wgsl:204-218) */
static float pd_convective_velocity(float heat_flux, float air_density, float temp_k, float hmix,
                                    const orc_pbl_options *o)
{
    if (heat_flux <= 0.0f) return 0.0f;
    const float d = pd_positive(air_density, 1.225f);
    const float t = pd_positive(temp_k, 300.0f);
    const float h = pd_positive(hmix, o->hmix_min_m);
    const float buoyancy = heat_flux / (d * PD_CPA) * PD_GA / t;
    if (buoyancy <= 0.0f || !pd_finite(buoyancy)) return 0.0f;
    return o_max(O_POW(buoyancy * h, 1.0f / 3.0f), 0.0f);
}

void orc_pbl_from_surface(uint64_t n, const float *const s0[ORC_SURF_COUNT], const float *const s1[ORC_SURF_COUNT],
                          float alpha, const orc_pbl_options *o, float *const out[7])
{
    for (uint64_t c = 0; c < n; ++c) {
        /* interpolate_array2_linear / lerp (io/temporal.rs:290-303); the
         * endpoints are exact copies (:134-139) */
        float f[ORC_SURF_COUNT];
        for (int k = 0; k < ORC_SURF_COUNT; ++k) {
            const float a = s0[k][c], b = s1[k][c];
            if (alpha <= 0.0f) f[k] = a;
            else if (alpha >= 1.0f) f[k] = b;
            else f[k] = fmaf(a, 1.0f - alpha, b * alpha);
        }
        const float ps = f[0], t2m = f[1], u10 = f[2], v10 = f[3], stress = f[4], sshf = f[5], ssr = f[6],
                    hmix_in = f[7], ustar_in = f[8], oli_in = f[9];
        /* main (pbl_diagnostics.wgsl:220-305) */
        const float temp_k = pd_positive(t2m, 300.0f);
        const float pressure = pd_positive(ps, 101325.0f);
        const float air_density = pressure / (PD_R_AIR * temp_k);
        const float estimated = pd_friction_velocity(stress, air_density, u10, v10, o->roughness_length_m,
                                                     o->wind_reference_height_m);
        float ustar;
        if (pd_finite(ustar_in) && ustar_in > 0.0f) ustar = o_max(ustar_in, PD_USTAR_MIN);
        else ustar = o_max(estimated, PD_USTAR_MIN);
        const float computed_ol = pd_obukhov_length(ustar, temp_k, air_density, sshf,
                                                    o->heat_flux_neutral_threshold_w_m2);
        float oli;
        if (pd_finite(oli_in) && fabsf(oli_in) >= PD_NEUTRAL_OLI) oli = oli_in;
        else oli = pd_inverse_obukhov(computed_ol);
        float hmix;
        if (pd_finite(hmix_in) && hmix_in > 0.0f) hmix = pd_clamp_hmix(hmix_in, o);
        else hmix = pd_clamp_hmix(o->fallback_mixing_height_m, o);
        const float safe_sshf = pd_fin(sshf, 0.0f);
        const float wstar = pd_convective_velocity(safe_sshf, air_density, temp_k, hmix, o);
        out[0][c] = ustar;
        out[1][c] = wstar;
        out[2][c] = hmix;
        out[3][c] = oli;
        out[4][c] = safe_sshf;
        out[5][c] = pd_nonneg(ssr, 0.0f);
        out[6][c] = pd_nonneg(stress, 0.0f);
    }
}

/* ---------------------------------------------------------------------- */
/* Concentration output: gpu/gridding.rs:194-219, docs/science/concentration-gridding.md */
/* ---------------------------------------------------------------------- */

/* dx_meters_at_latitude (coords/mod.rs:156-158), f32 with R_EARTH = 6.371e6
 * (lib.rs:22) and PI180 = pi/180 (lib.rs:32) */
static float out_dx_meters(float lat_deg)
{
    const float r_earth = 6.371e6f, pi180 = 3.14159265358979323846f / 180.0f;
    return r_earth * cosf(lat_deg * pi180) * pi180;
}

void orc_output_volumes(double ylat0, double dx, double dy, const float *outheights, uint32_t ny, uint32_t nz,
                        float *vol)
{
    const double dy_m = (double)out_dx_meters(0.0f); /* R PI180: cos 0 = 1 */
    for (uint32_t iy = 0; iy < ny; ++iy) {
        const float lat = (float)(ylat0 + ((double)iy + 0.5) * dy);
        const double area = (dx * (double)out_dx_meters(lat)) * (dy * dy_m);
        for (uint32_t k = 0; k < nz; ++k) {
            const double dz = (double)outheights[k] - (k ? (double)outheights[k - 1] : 0.0);
            vol[(uint64_t)iy * nz + k] = (float)(area * dz);
        }
    }
}

void orc_output_finalize(uint64_t ncell, const uint64_t *acc_count, const uint64_t *acc_mass, uint32_t n_samples,
                         float mass_scale, const float *vol, uint32_t ny, uint32_t nz, uint32_t *count, float *mass,
                         float *conc)
{
    for (uint64_t c = 0; c < ncell; ++c) {
        count[c] = acc_count[c] > 0xffffffffull ? 0xffffffffu : (uint32_t)acc_count[c];
        const float m = (float)((double)acc_mass[c] / (double)n_samples) / mass_scale;
        mass[c] = m;
        if (vol && conc) {
            const uint32_t k = (uint32_t)(c % nz), iy = (uint32_t)((c / nz) % ny);
            conc[c] = m / vol[(uint64_t)iy * nz + k];
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Hybrid sigma-pressure to height: io/vertical_transform.rs                */
/* ---------------------------------------------------------------------- */

/* virtual_temperature_k (vertical_transform.rs:272-292) */
static int vt_virtual_temperature(float t, float q, float *tv)
{
    if (!(t == t && fabsf(t) <= FLT_MAX) || t <= 0.0f || !(q == q && fabsf(q) <= FLT_MAX)) return 0;
    *tv = t * (1.0f + 0.608f * o_max(q, 0.0f));
    return 1;
}

/* hypsometric_layer_thickness_m (vertical_transform.rs:294-313) */
static float vt_thickness(float p_prev, float p_cur, float tv_prev, float tv_cur)
{
    const float scale = 287.05f / 9.81f; /* R_AIR / GA */
    const float pr_ln = O_LOG(p_prev / p_cur);
    const float dtv = tv_cur - tv_prev;
    const float tvr_ln = O_LOG(tv_cur / tv_prev);
    if (fabsf(dtv) > 0.2f && fabsf(tvr_ln) > FLT_EPSILON) return scale * pr_ln * dtv / tvr_ln;
    return scale * pr_ln * tv_cur;
}

int orc_hybrid_to_height(uint32_t nx, uint32_t ny, uint32_t nz, const float *ps, const float *t, const float *q,
                         const float *ak, const float *bk, float *pressure, float *height, float *mean,
                         uint32_t err_ijk[3])
{
    /* compute_hybrid_level_pressure_pa */
    for (uint32_t i = 0; i < nx; ++i)
        for (uint32_t j = 0; j < ny; ++j)
            for (uint32_t k = 0; k < nz; ++k) {
                const uint64_t c = ((uint64_t)i * ny + j) * nz + k;
                const float p = ak[k] + bk[k] * ps[(uint64_t)i * ny + j];
                if (!(p == p && fabsf(p) <= FLT_MAX) || p <= 0.0f) {
                    err_ijk[0] = i; err_ijk[1] = j; err_ijk[2] = k;
                    return 1;
                }
                pressure[c] = p;
            }
    for (uint32_t i = 0; i < nx; ++i)
        for (uint32_t j = 0; j < ny; ++j) {
            const uint64_t base = ((uint64_t)i * ny + j) * nz;
            float p_prev = pressure[base], tv_prev;
            height[base] = 0.0f;
            if (!vt_virtual_temperature(t[base], q ? q[base] : 0.0f, &tv_prev)) {
                err_ijk[0] = i; err_ijk[1] = j; err_ijk[2] = 0;
                return 3;
            }
            for (uint32_t k = 1; k < nz; ++k) {
                const float p = pressure[base + k];
                if (p >= p_prev) {
                    err_ijk[0] = i; err_ijk[1] = j; err_ijk[2] = k;
                    return 2;
                }
                float tv;
                if (!vt_virtual_temperature(t[base + k], q ? q[base + k] : 0.0f, &tv)) {
                    err_ijk[0] = i; err_ijk[1] = j; err_ijk[2] = k;
                    return 3;
                }
                const float z_prev = height[base + k - 1];
                const float z = z_prev + vt_thickness(p_prev, p, tv_prev, tv);
                if (!(z == z && fabsf(z) <= FLT_MAX) || z <= z_prev) {
                    err_ijk[0] = i; err_ijk[1] = j; err_ijk[2] = k;
                    return 4;
                }
                height[base + k] = z;
                p_prev = p;
                tv_prev = tv;
            }
        }
    /* horizontal_mean_height_profile_m */
    const float count = (float)((uint64_t)nx * ny);
    for (uint32_t k = 0; k < nz; ++k) {
        float s = 0.0f;
        for (uint32_t i = 0; i < nx; ++i)
            for (uint32_t j = 0; j < ny; ++j) s += height[((uint64_t)i * ny + j) * nz + k];
        mean[k] = s / count;
    }
    return 0;
}
