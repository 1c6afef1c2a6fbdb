// fp_adj.hip - differentiable runs: a forward run of point sources through
// the timestep and its hand-written adjoint (fp_adj_* in
// include/flexpart_hip.h).
//
// Forward: the emission q_s and release position r_s of every source set the
// initial particles; the wind perturbation sum_k c_k (Bu_k, Bv_k) is added to
// the met; the production step kernel advances the particles; at the
// observation times every particle adds m W(X, node) to the receptors on its
// 8 surrounding lattice nodes (cloud-in-cell weights, f64 sums).  States are
// checkpointed every K steps.
//
// Backward: for each K-step segment, last to first, the segment is recomputed
// from its checkpoint (Philox is counter-based, so the draws replay exactly)
// keeping the state at the start of every step, then the adjoint kernel walks
// the steps back.  Per particle and step it recomputes the step from the
// saved state with the shared device functions (bit-identical values),
// keeping what the reverse sweep needs: the wind Jacobians at the predictor
// and corrector points, the clamp masks, the Hanna profiles and their height
// derivatives (forward-mode Dz), the Langevin coefficients and the
// reflection signs of every sub-step.  It then maps the adjoint state
// (x, y, z, u', v', w', m) of the step's output to its input and scatters the
// adjoint of the sampled horizontal winds onto two 2-D fields whose dot
// products with the basis give dL/dc_k.  Discontinuities (cell lookups,
// regime and branch switches, deposition thresholds) contribute nothing: the
// result is the pathwise derivative.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fp_physics.hpp"

static_assert(sizeof(fp_adj_desc) == 56, "fp_adj_desc layout is part of the ABI");

namespace fp {
namespace {

// Fields a step reads and writes for one species: what checkpoints and the
// per-step tape keep.
constexpr uint32_t kStateFields[] = {F_POS_X, F_POS_Y, F_POS_Z, F_CELL_X, F_CELL_Y, F_FLAGS,
                                     F_MASS0, F_TURB_U, F_TURB_V, F_TURB_W, F_CBT};
constexpr uint32_t kLam = 7;  // adjoint state: x, y, z, u', v', w', m
enum : uint32_t { L_X = 0, L_Y, L_Z, L_TU, L_TV, L_TW, L_M };

struct ObsArgs {
    const int32_t *node_rec;  // lattice node -> receptor index or -1
    uint32_t nx, ny, nz;      // lattice: met columns x obs levels
    float dz;
};

struct AdjArgs {
    StepArgs a;         // the step (a.s = the state at its start)
    float *lam;         // kLam x stride
    uint64_t stride;
    double *au, *av;    // adjoint of the horizontal wind perturbation, nx * ny
};

// ---------------------------------------------------------------- wind with gradient

// Derivatives of the sampled u, v, w (sample_wind) with respect to x, y
// (grid units) and z (m), and the horizontal cell and weights of the sample
// (for the perturbation adjoint).
struct WindGrad {
    float ux, uy, uz, vx, vy, vz, wx, wy, wz;
    uint32_t i0, i1, j0, j1;
    float tx, ty;
};

__device__ __forceinline__ void level_grad(float z, const float *lh, uint32_t nz, uint32_t heights, float &lev,
                                           float &dlev)
{
    lev = level_of(z, lh, nz, heights);
    if (!heights) { dlev = 1.0f; return; }
    const float h0 = lh[0];
    if (z <= h0) { dlev = h0 > 0.0f ? 1.0f / h0 : 0.0f; return; }
    uint32_t lo = 1, hi = nz;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (z <= lh[mid]) hi = mid; else lo = mid + 1;
    }
    if (lo >= nz) { dlev = 0.0f; return; }
    const float den = lh[lo] - lh[lo - 1];
    dlev = den > 0.0f ? 1.0f / den : 0.0f;
}

__device__ __forceinline__ WindGrad sample_wind_grad(const StepArgs &a, const float *lh, float alpha, float x,
                                                     float y, float z)
{
    WindGrad g;
    // sum over the 8 corners of d(weight)/d(coordinate) * value, corner by
    // corner (the step's own sample_wind read the same lines)
    const Axis ax = axis_of(x, a.nx, a.nxm);
    const Axis ay = axis_of(y, a.ny, a.nym);
    float lev, dlev;
    level_grad(z, lh, a.nz, a.has_heights, lev, dlev);
    const Axis az = a.has_heights ? axis_of(lev, a.nz, a.nzm) : axis_of<true>(lev, a.nz, a.nzm);
    const uint32_t ny = a.ny, nz = a.nz;
    const float ux = 1.0f - ax.t, uy = 1.0f - ay.t, uz = 1.0f - az.t, ua = 1.0f - alpha;
    float gx[3] = {0, 0, 0}, gy[3] = {0, 0, 0}, gz[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t ix = (k & 1) ? ax.hi : ax.lo, iy = (k & 2) ? ay.hi : ay.lo, iz = (k & 4) ? az.hi : az.lo;
        const uint32_t idx = (ix * ny + iy) * nz + iz;
        const float4 p = load_at(a.met, idx << 5);
        const float4 q = load_at(a.met, idx << 5, 16u);
        const float wx = (k & 1) ? ax.t : ux, wy = (k & 2) ? ay.t : uy, wz = (k & 4) ? az.t : uz;
        const float sx = (k & 1) ? 1.0f : -1.0f, sy = (k & 2) ? 1.0f : -1.0f, sz = (k & 4) ? 1.0f : -1.0f;
        // time-blended corner value of u, v, w
        const float cu = ua * p.x + alpha * p.w, cv = ua * p.y + alpha * q.x, cw = ua * p.z + alpha * q.y;
        const float dx = sx * wy * wz, dy = wx * sy * wz, dz = wx * wy * sz;
        gx[0] += dx * cu; gx[1] += dx * cv; gx[2] += dx * cw;
        gy[0] += dy * cu; gy[1] += dy * cv; gy[2] += dy * cw;
        gz[0] += dz * cu; gz[1] += dz * cv; gz[2] += dz * cw;
    }
    // chain to x, y, z: zero where the coordinate is clamped or the axis is a single point
    const float fx = (x > 0.0f && x < (float)(a.nx - 1u)) ? 1.0f : 0.0f;
    const float fy = (y > 0.0f && y < (float)(a.ny - 1u)) ? 1.0f : 0.0f;
    const float fz = (lev > 0.0f && lev < (float)(a.nz - 1u)) ? dlev : 0.0f;
    g.ux = fx * gx[0]; g.vx = fx * gx[1]; g.wx = fx * gx[2];
    g.uy = fy * gy[0]; g.vy = fy * gy[1]; g.wy = fy * gy[2];
    g.uz = fz * gz[0]; g.vz = fz * gz[1]; g.wz = fz * gz[2];
    g.i0 = ax.lo; g.i1 = ax.hi; g.j0 = ay.lo; g.j1 = ay.hi;
    g.tx = ax.t; g.ty = ay.t;
    return g;
}

// Wave-level sum of the (u, v) perturbation adjoints of one wind sample over
// lanes that share its horizontal cell, then 8 f64 atomics per distinct cell.
__device__ __forceinline__ void scatter_wind_adjoint(const AdjArgs &g, bool valid, const WindGrad &s, float lu,
                                                     float lv)
{
    const uint32_t ny = g.a.ny;
    const uint32_t key = valid ? s.i0 * ny + s.j0 : 0xffffffffu;
    const float w00 = (1.0f - s.tx) * (1.0f - s.ty), w10 = s.tx * (1.0f - s.ty);
    const float w01 = (1.0f - s.tx) * s.ty, w11 = s.tx * s.ty;
    uint64_t pending = __ballot(valid && (lu != 0.0f || lv != 0.0f));
    const uint32_t lane = __lane_id();
    while (pending) {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)pending) - 1u;
        const uint32_t lk = __shfl(key, (int)leader);
        const bool mine = valid && key == lk;
        const uint64_t same = __ballot(mine);
        float p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (mine) {
            p[0] = lu * w00; p[1] = lu * w10; p[2] = lu * w01; p[3] = lu * w11;
            p[4] = lv * w00; p[5] = lv * w10; p[6] = lv * w01; p[7] = lv * w11;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) p[k] += __shfl_xor(p[k], off);
        if (lane == leader) {
            const uint32_t c[4] = {s.i0 * ny + s.j0, s.i1 * ny + s.j0, s.i0 * ny + s.j1, s.i1 * ny + s.j1};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (p[k] != 0.0f) atomicAdd(&g.au[c[k]], (double)p[k]);
                if (p[4 + k] != 0.0f) atomicAdd(&g.av[c[k]], (double)p[4 + k]);
            }
        }
        pending &= ~same;
    }
}

// ---------------------------------------------------------------- adjoint step

// Capped at 128 VGPRs (4 wavefronts per SIMD): the uncapped kernel takes 147
// and runs 3; the cap spills 18 registers to scratch and still runs the
// reverse sweep 2 % faster.
template <int NSUB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) adj_step_kernel(AdjArgs g)
{
    __shared__ float lh[kLevelTableMax];
    const StepArgs &a = g.a;
    for (uint32_t k = threadIdx.x; k < level_table_words(a.nz); k += 256) lh[k] = a.level_heights[k];
    load_math_tab<256>();
    __syncthreads();
    const uint32_t nsub = NSUB >= 0 ? (uint32_t)NSUB : a.n_substeps;
    const Soa s = a.s;
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    bool live = i < a.n;
    if (live) live = (s.u(F_FLAGS)[i] & FLAG_ACTIVE) != 0u;
    WindGrad s0{}, s1{};
    float lu0 = 0.0f, lv0 = 0.0f, lu1 = 0.0f, lv1 = 0.0f;
    if (live) {
        const int32_t cx = s.i(F_CELL_X)[i], cy = s.i(F_CELL_Y)[i];
        const float px = s.f(F_POS_X)[i], py = s.f(F_POS_Y)[i], pz = s.f(F_POS_Z)[i];
        const float tu = s.f(F_TURB_U)[i], tv = s.f(F_TURB_V)[i], tw = s.f(F_TURB_W)[i];
        const float alpha = a.alpha;

        // ---- forward recomputation (the step kernel's operations)
        const float nxm = a.nxm, nym = a.nym;
        const float x0 = (float)cx + px, y0 = (float)cy + py, z0 = pz;
        const Wind sw0 = sample_wind(a, lh, alpha, x0, y0, z0);
        const float u0 = sw0.u + tu, v0 = sw0.v + tv, w0 = sw0.w;
        const float ex = x0 + (u0 * a.x_scale) * a.dt_adv;
        const float ey = y0 + (v0 * a.y_scale) * a.dt_adv;
        const float ez = z0 + (w0 * a.z_scale) * a.dt_adv;
        const bool mqx = ex > 0.0f && ex < nxm, mqy = ey > 0.0f && ey < nym, mqz = ez > 0.0f && ez < a.zmax;
        const float qx = clampf(ex, 0.0f, nxm), qy = clampf(ey, 0.0f, nym), qz = clampf(ez, 0.0f, a.zmax);
        const Wind sw1 = sample_wind(a, lh, alpha, qx, qy, qz);
        const float u1 = sw1.u + tu, v1 = sw1.v + tv, w1 = sw1.w;
        const float cu = 0.5f * (u0 + u1), cv = 0.5f * (v0 + v1), cw = 0.5f * (w0 + w1);
        const float fx = x0 + (cu * a.x_scale) * a.dt_adv;
        const float fy = y0 + (cv * a.y_scale) * a.dt_adv;
        const float fz = z0 + (cw * a.z_scale) * a.dt_adv;
        const bool m1x = fx > 0.0f && fx < nxm, m1y = fy > 0.0f && fy < nym, m1z = fz > 0.0f && fz < a.zmax;
        const float x1 = clampf(fx, 0.0f, nxm), y1 = clampf(fy, 0.0f, nym);
        const float z1 = clampf(fz, 0.0f, a.zmax);
        int32_t ncx, ncy;
        float npx, npy;
        split_cell(x1, a.nx, nxm, ncx, npx);
        split_cell(y1, a.ny, nym, ncy, npy);
        const float hx = (float)ncx + npx, hy = (float)ncy + npy;
        const uint32_t pi = (uint32_t)floorf(clampf(hx, 0.0f, nxm));
        const uint32_t pj = (uint32_t)floorf(clampf(hy, 0.0f, nym));
        const uint32_t pidx = pi * a.ny + pj;
        const TurbT<Dz> t = hanna_full(a.pbl[2u * pidx], a.pbl[2u * pidx + 1u], Dz{z1, 1.0f}, a.log_1em3,
                                           a.lean_div != 0u);

        const U4 ctr{a.c0, a.c1, a.c2, a.c3};
        const uint64_t cbase = i * a.blocks_per_particle;
        const U4 b0 = philox10(ctr_add(ctr.x, ctr.y, ctr.z, ctr.w, cbase), a.key0, a.key1);
        U4 b1 = U4{0u, 0u, 0u, 0u};
        if (a.blocks_per_particle > 1u)
            b1 = philox10(ctr_add(ctr.x, ctr.y, ctr.z, ctr.w, cbase + 1u), a.key0, a.key1);
        float eu, ev, ew0, ew1, ew2, ew3;
        gauss_pair(b0.x, b0.y, eu, ev);
        gauss_pair(b0.z, b0.w, ew0, ew1);
        gauss_pair(b1.x, b1.y, ew2, ew3);

        float cu_prev, cv_prev;  // d u'/d u, d v'/d v (the retained fractions)
        Dz tun, tvn;
        if (t.lean) {
            tun = ou_horizontal<true>(Dz{tu, 0.0f}, t.sigu, t.tlu, a.dt, eu, &cu_prev);
            tvn = ou_horizontal<true>(Dz{tv, 0.0f}, t.sigv, t.tlv, a.dt, ev, &cv_prev);
        } else {
            tun = ou_horizontal(Dz{tu, 0.0f}, t.sigu, t.tlu, a.dt, eu, &cu_prev);
            tvn = ou_horizontal(Dz{tv, 0.0f}, t.sigv, t.tlv, a.dt, ev, &cv_prev);
        }

        float A[4], G[4], S[4];
        bool K[4];
        float zf = z1;
        if (nsub == 0u) {
            const Dz twn = t.lean ? ou_vertical<true>(Dz{tw, 0.0f}, t.sigw, t.dsigwdz, t.tlw, a.dt, a.rho_grad, ew0, &A[0])
                                  : ou_vertical(Dz{tw, 0.0f}, t.sigw, t.dsigwdz, t.tlw, a.dt, a.rho_grad, ew0, &A[0]);
            G[0] = twn.d;
        } else {
            const float dts = a.dt_sub;  // a.dt / n_substeps
            const float zmin = a.min_height, h = t.h;
            Dz sigw = t.sigw, dsw = t.dsigwdz, tlw = t.tlw;
            float z = z1, w = tw;
#pragma unroll
            for (uint32_t sub = 0; sub < (NSUB > 0 ? (uint32_t)NSUB : 4u); ++sub) {
                if (NSUB < 0 && sub >= nsub) break;
                const float eta = sub == 0u ? ew0 : (sub == 1u ? ew1 : (sub == 2u ? ew2 : ew3));
                Dz wn;
                if (t.lean) wn = ou_vertical<true>(Dz{w, 0.0f}, sigw, dsw, tlw, dts, a.rho_grad, eta, &A[sub]);
                else wn = ou_vertical(Dz{w, 0.0f}, sigw, dsw, tlw, dts, a.rho_grad, eta, &A[sub]);
                G[sub] = wn.d;
                float zz = z + wn.v * dts;
                float sg = 1.0f;
                bool kap = true;
                if (h > zmin) {  // reflect() with the sign of the velocity tracked
                    for (int bounce = 0; bounce < 10; ++bounce) {
                        if (zz < zmin) { zz = 2.0f * zmin - zz; sg = -sg; }
                        else if (zz > h) { zz = 2.0f * h - zz; sg = -sg; }
                        else break;
                    }
                    kap = zz >= zmin && zz <= h;
                    zz = clampf(zz, zmin, h);
                }
                z = zz;
                w = sg * wn.v;
                S[sub] = sg;
                K[sub] = kap;
                if (sub + 1u < nsub) vertical_profile(t, Dz{z, 1.0f}, a.log_1em3, sigw, dsw, tlw);
            }
            zf = z;
        }

        // survival factor of the mass (dry and wet deposition)
        float fm = 1.0f;
        if (a.dry_mode != FP_DRY_OFF) {
            const float href = fmaxf(a.dry_href, 0.1f);
            if (!(zf >= 2.0f * href)) {
                const float vd = a.dry_mode == FP_DRY_PER_SLOT ? a.dry_vd_slots[i] : a.dry_vd;
                const float ex2 = -fmaxf(vd, 0.0f) * fabsf(a.dt) / (2.0f * href);
                fm = fm * clampf(fpm_expf(ex2), 0.0f, 1.0f);
            }
        }
        if (a.wet_mode != FP_WET_OFF) {
            float lam, frac;
            if (a.wet_mode == FP_WET_PER_SLOT) {
                lam = a.wet_lambda_slots[i];
                frac = a.wet_fraction_slots[i];
            } else if (a.wet_mode == FP_WET_PRECIP_FIELD) {
                const float pr = pos_or_zero(a.precip[pidx]);
                lam = 0.0f;
                if (pr > 0.0f) lam = pos_or_zero(pos_or_zero(a.wet_a) * fpm_powf(pr, pos_or_zero(a.wet_b)));
                frac = a.wet_fraction;
            } else {
                lam = a.wet_lambda;
                frac = a.wet_fraction;
            }
            const float l = fmaxf(lam, 0.0f), dt = fabsf(a.dt), fr = clampf(frac, 0.0f, 1.0f);
            float prob = 0.0f;
            if (!(l <= 0.0f || dt <= 0.0f || fr <= 0.0f)) prob = clampf((1.0f - fpm_expf(-l * dt)) * fr, 0.0f, 1.0f);
            fm = fm * (1.0f - prob);
        }

        // ---- reverse sweep
        float *L = g.lam;
        const uint64_t st = g.stride;
        float lx = L[L_X * st + i], ly = L[L_Y * st + i], lz = L[L_Z * st + i];
        float ltu = L[L_TU * st + i], ltv = L[L_TV * st + i], ltw = L[L_TW * st + i];
        const float lm = L[L_M * st + i] * fm;

        if (nsub == 0u) {
            lz = lz + G[0] * ltw;
            ltw = A[0] * ltw;
        } else {
            const float dts = a.dt_sub;  // a.dt / n_substeps
#pragma unroll
            for (int sub = (NSUB > 0 ? NSUB : 4) - 1; sub >= 0; --sub) {
                if (NSUB < 0 && (uint32_t)sub >= nsub) continue;
                const float lzp = S[sub] * (K[sub] ? lz : 0.0f);
                const float lt = S[sub] * ltw + dts * lzp;
                lz = lzp + G[sub] * lt;
                ltw = A[sub] * lt;
            }
        }
        // horizontal Langevin: u'(u, z1), v'(v, z1)
        lz = lz + tun.d * ltu + tvn.d * ltv;
        float ltu_in = cu_prev * ltu, ltv_in = cv_prev * ltv;

        // advection: x1 = clamp(x0 + c_u s dt), c = (S0 + S1)/2 + (turb, 0)
        const float lcu = m1x ? lx * a.x_scale * a.dt_adv : 0.0f;
        const float lcv = m1y ? ly * a.y_scale * a.dt_adv : 0.0f;
        const float lcw = m1z ? lz * a.z_scale * a.dt_adv : 0.0f;
        float lx0 = m1x ? lx : 0.0f, ly0 = m1y ? ly : 0.0f, lz0 = m1z ? lz : 0.0f;
        lu1 = 0.5f * lcu; lv1 = 0.5f * lcv;
        const float lw1 = 0.5f * lcw;
        lu0 = 0.5f * lcu; lv0 = 0.5f * lcv;
        float lw0 = 0.5f * lcw;
        ltu_in += lu1;
        ltv_in += lv1;
        // corrector sample at q (its Jacobian re-derived from the corners)
        s1 = sample_wind_grad(a, lh, alpha, qx, qy, qz);
        const float lqx = lu1 * s1.ux + lv1 * s1.vx + lw1 * s1.wx;
        const float lqy = lu1 * s1.uy + lv1 * s1.vy + lw1 * s1.wy;
        const float lqz = lu1 * s1.uz + lv1 * s1.vz + lw1 * s1.wz;
        if (mqx) { lx0 += lqx; lu0 += lqx * a.x_scale * a.dt_adv; }
        if (mqy) { ly0 += lqy; lv0 += lqy * a.y_scale * a.dt_adv; }
        if (mqz) { lz0 += lqz; lw0 += lqz * a.z_scale * a.dt_adv; }
        ltu_in += lu0;
        ltv_in += lv0;
        // predictor sample at x0
        s0 = sample_wind_grad(a, lh, alpha, x0, y0, z0);
        lx0 += lu0 * s0.ux + lv0 * s0.vx + lw0 * s0.wx;
        ly0 += lu0 * s0.uy + lv0 * s0.vy + lw0 * s0.wy;
        lz0 += lu0 * s0.uz + lv0 * s0.vz + lw0 * s0.wz;

        L[L_X * st + i] = lx0;
        L[L_Y * st + i] = ly0;
        L[L_Z * st + i] = lz0;
        L[L_TU * st + i] = ltu_in;
        L[L_TV * st + i] = ltv_in;
        L[L_TW * st + i] = ltw;
        L[L_M * st + i] = lm;
    }
    if (g.au) {
        scatter_wind_adjoint(g, live, s0, lu0, lv0);
        scatter_wind_adjoint(g, live, s1, lu1, lv1);
    }
}

// ---------------------------------------------------------------- observations

struct Cic {
    uint32_t i0, j0, l0;
    float fx, fy, fz;
    float sx, sy, sz;  // d fx/dx, d fy/dy, d fz/dz (0 where clamped)
};

__device__ __forceinline__ Cic cic_of(const ObsArgs &o, float x, float y, float z)
{
    Cic c;
    const float xm = (float)(o.nx - 1u), ym = (float)(o.ny - 1u), zm = (float)(o.nz - 1u) * o.dz;
    const float xc = clampf(x, 0.0f, xm), yc = clampf(y, 0.0f, ym);
    const float zc = clampf(z, 0.0f, zm) / o.dz;
    c.i0 = min((uint32_t)floorf(xc), o.nx - 2u);
    c.j0 = min((uint32_t)floorf(yc), o.ny - 2u);
    c.l0 = min((uint32_t)floorf(zc), o.nz - 2u);
    c.fx = xc - (float)c.i0;
    c.fy = yc - (float)c.j0;
    c.fz = zc - (float)c.l0;
    c.sx = (x > 0.0f && x < xm) ? 1.0f : 0.0f;
    c.sy = (y > 0.0f && y < ym) ? 1.0f : 0.0f;
    c.sz = (z > 0.0f && z < zm) ? 1.0f / o.dz : 0.0f;
    return c;
}

__device__ __forceinline__ uint32_t cic_node(const ObsArgs &o, const Cic &c, int k)
{
    return ((c.i0 + (uint32_t)(k & 1)) * o.ny + c.j0 + (uint32_t)((k >> 1) & 1)) * o.nz + c.l0 +
           (uint32_t)(k >> 2);
}

__device__ __forceinline__ float cic_weight(const Cic &c, int k)
{
    const float wx = (k & 1) ? c.fx : 1.0f - c.fx;
    const float wy = (k & 2) ? c.fy : 1.0f - c.fy;
    const float wz = (k & 4) ? c.fz : 1.0f - c.fz;
    return wx * wy * wz;
}

__global__ void __launch_bounds__(256) obs_deposit_kernel(Soa s, uint32_t n, ObsArgs o, double *rec_val)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    bool live = i < n;
    if (live) live = (s.u(F_FLAGS)[i] & FLAG_ACTIVE) != 0u;
    Cic c{};
    float m = 0.0f;
    if (live) {
        c = cic_of(o, (float)s.i(F_CELL_X)[i] + s.f(F_POS_X)[i], (float)s.i(F_CELL_Y)[i] + s.f(F_POS_Y)[i],
                   s.f(F_POS_Z)[i]);
        m = s.f(F_MASS0)[i];
    }
    const uint32_t lane = __lane_id();
#pragma unroll 1
    for (int k = 0; k < 8; ++k) {
        const int32_t rec = live ? o.node_rec[cic_node(o, c, k)] : -1;
        const bool valid = rec >= 0;
        const float val = valid ? m * cic_weight(c, k) : 0.0f;
        uint64_t pending = __ballot(valid);
        while (pending) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)pending) - 1u;
            const int32_t lr = __shfl(rec, (int)leader);
            const bool mine = valid && rec == lr;
            const uint64_t same = __ballot(mine);
            double v = mine ? (double)val : 0.0;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
            if (lane == leader) atomicAdd(&rec_val[lr], v);
            pending &= ~same;
        }
    }
}

// adjoint of the observation at one time: lambda(state) += dC/dstate^T rec_lam
__global__ void __launch_bounds__(256) obs_adjoint_kernel(Soa s, uint32_t n, ObsArgs o, const double *rec_lam,
                                                          float *lam, uint64_t stride)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= n || !(s.u(F_FLAGS)[i] & FLAG_ACTIVE)) return;
    const Cic c = cic_of(o, (float)s.i(F_CELL_X)[i] + s.f(F_POS_X)[i], (float)s.i(F_CELL_Y)[i] + s.f(F_POS_Y)[i],
                         s.f(F_POS_Z)[i]);
    const float m = s.f(F_MASS0)[i];
    float gm = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int32_t rec = o.node_rec[cic_node(o, c, k)];
        if (rec < 0) continue;
        const float lc = (float)rec_lam[rec];
        const float wx = (k & 1) ? c.fx : 1.0f - c.fx, dwx = (k & 1) ? 1.0f : -1.0f;
        const float wy = (k & 2) ? c.fy : 1.0f - c.fy, dwy = (k & 2) ? 1.0f : -1.0f;
        const float wz = (k & 4) ? c.fz : 1.0f - c.fz, dwz = (k & 4) ? 1.0f : -1.0f;
        gm += wx * wy * wz * lc;
        gx += dwx * wy * wz * lc;
        gy += wx * dwy * wz * lc;
        gz += wx * wy * dwz * lc;
    }
    lam[L_M * stride + i] += gm;
    lam[L_X * stride + i] += m * c.sx * gx;
    lam[L_Y * stride + i] += m * c.sy * gy;
    lam[L_Z * stride + i] += m * c.sz * gz;
}

// ---------------------------------------------------------------- sources, met, reductions

__global__ void __launch_bounds__(256) init_sources_kernel(Soa s, uint32_t n, const uint32_t *src_of,
                                                           const uint32_t *counts, const float *q,
                                                           const float *r, uint32_t nx, uint32_t ny)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = src_of[i];
    for (uint32_t f = 0; f < F_COUNT; ++f) s.u(f)[i] = 0u;
    int32_t cx, cy;
    float px, py;
    split_cell(r[3u * k], nx, cx, px);
    split_cell(r[3u * k + 1u], ny, cy, py);
    s.i(F_CELL_X)[i] = cx;
    s.i(F_CELL_Y)[i] = cy;
    s.f(F_POS_X)[i] = px;
    s.f(F_POS_Y)[i] = py;
    s.f(F_POS_Z)[i] = r[3u * k + 2u];
    s.f(F_MASS0)[i] = q[k] / (float)counts[k];
    s.u(F_FLAGS)[i] = FLAG_ACTIVE;
    s.i(F_CBT)[i] = 1;
}

// perturbation fields p(i,j) = sum_k c_k B_k(i,j) for u and v
__global__ void __launch_bounds__(256) perturb_kernel(const float *bu, const float *bv, const float *coef,
                                                      uint32_t K, uint32_t nxy, float *pu, float *pv)
{
    const uint32_t ij = blockIdx.x * 256u + threadIdx.x;
    if (ij >= nxy) return;
    float su = 0.0f, sv = 0.0f;
    for (uint32_t k = 0; k < K; ++k) {
        const float c = coef[k];
        su = su + c * bu[(uint64_t)k * nxy + ij];
        sv = sv + c * bv[(uint64_t)k * nxy + ij];
    }
    pu[ij] = su;
    pv[ij] = sv;
}

// packed bracket b: {u_b, v_b, w_b, u_{b+1}} {v_{b+1}, w_{b+1}, 0, 0} with the perturbation added
__global__ void __launch_bounds__(256) pack_bracket_kernel(const float *base, uint64_t npts, uint32_t nz,
                                                           uint32_t b, const float *pu, const float *pv,
                                                           float4 *met)
{
    const uint64_t p = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (p >= npts) return;
    const uint32_t ij = (uint32_t)(p / nz);
    const float *l0 = base + (uint64_t)b * 3u * npts;
    const float *l1 = l0 + 3u * npts;
    const float du = pu ? pu[ij] : 0.0f, dv = pv ? pv[ij] : 0.0f;
    met[2u * p] = make_float4(l0[p] + du, l0[npts + p] + dv, l0[2u * npts + p], l1[p] + du);
    met[2u * p + 1u] = make_float4(l1[npts + p] + dv, l1[2u * npts + p], 0.0f, 0.0f);
}

// gradient of every source: one block per source over its particle range
__global__ void __launch_bounds__(256) source_grad_kernel(const float *lam, uint64_t stride, const uint64_t *offs,
                                                          const uint32_t *counts, const float *r, uint32_t nx,
                                                          uint32_t ny, double *gq, double *gr)
{
    const uint32_t k = blockIdx.x;
    const uint64_t a = offs[k], n = counts[k];
    double sm = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    for (uint64_t p = threadIdx.x; p < n; p += 256u) {
        sm += (double)lam[L_M * stride + a + p];
        sx += (double)lam[L_X * stride + a + p];
        sy += (double)lam[L_Y * stride + a + p];
        sz += (double)lam[L_Z * stride + a + p];
    }
    __shared__ double red[4][256];
    red[0][threadIdx.x] = sm; red[1][threadIdx.x] = sx; red[2][threadIdx.x] = sy; red[3][threadIdx.x] = sz;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int f = 0; f < 4; ++f) red[f][threadIdx.x] += red[f][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float rx = r[3u * k], ry = r[3u * k + 1u];
        gq[k] = red[0][0] / (double)n;
        gr[3u * k] = (rx > 0.0f && rx < (float)(nx - 1u)) ? red[1][0] : 0.0;
        gr[3u * k + 1u] = (ry > 0.0f && ry < (float)(ny - 1u)) ? red[2][0] : 0.0;
        gr[3u * k + 2u] = red[3][0];
    }
}

// dL/dc_k = <Au, Bu_k> + <Av, Bv_k>: one block per mode
__global__ void __launch_bounds__(256) coeff_grad_kernel(const double *au, const double *av, const float *bu,
                                                         const float *bv, uint32_t nxy, double *gc)
{
    const uint32_t k = blockIdx.x;
    double sum = 0.0;
    for (uint32_t ij = threadIdx.x; ij < nxy; ij += 256u)
        sum += au[ij] * (double)bu[(uint64_t)k * nxy + ij] + av[ij] * (double)bv[(uint64_t)k * nxy + ij];
    __shared__ double red[256];
    red[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) gc[k] = red[0];
}

__global__ void __launch_bounds__(256) fill_f32_kernel(float *p, uint64_t n, float v)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = v;
}

// zero rows 0..nf-1 of a quantity-major buffer over its first n columns
__global__ void __launch_bounds__(256) zero_rows_kernel(float *p, uint64_t stride, uint32_t nf, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= n * nf) return;
    p[(i / n) * stride + i % n] = 0.0f;
}

__global__ void __launch_bounds__(256) fill_f64_kernel(double *p, uint64_t n, double v)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = v;
}

// the state fields of survivor j <- slot map[j] (replay of a saved compaction)
__global__ void __launch_bounds__(256) gather_state_kernel(Soa in, Soa out, const uint32_t *map, uint32_t n_out)
{
    const uint64_t j = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (j >= n_out) return;
    const uint32_t src = map[j];
#pragma unroll
    for (uint32_t f : kStateFields) out.u(f)[j] = in.u(f)[src];
}

// adjoint of a compaction: lam_pre[map[j]] = lam_post[j] (lam_pre zeroed
// before: removed particles carry no adjoint)
__global__ void __launch_bounds__(256) uncompact_kernel(const float *post, float *pre, const uint32_t *map,
                                                        uint32_t n_post, uint64_t stride, uint32_t nf)
{
    const uint64_t j = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (j >= n_post) return;
    const uint32_t p = map[j];
    for (uint32_t f = 0; f < nf; ++f) pre[f * stride + p] = post[f * stride + j];
}

// Survival factor of the mass over a step (dry and wet deposition), from the
// particle's state after the step: the final height and the PBL cell of the
// post-advection position (the step kernel's deposition, same operations).
__device__ __forceinline__ float survival_factor(const StepArgs &a, const Soa &s, uint64_t j)
{
    float fm = 1.0f;
    const float zf = s.f(F_POS_Z)[j];
    if (a.dry_mode != FP_DRY_OFF) {
        const float href = fmaxf(a.dry_href, 0.1f);
        if (!(zf >= 2.0f * href)) {
            const float ex = -fmaxf(a.dry_vd, 0.0f) * fabsf(a.dt) / (2.0f * href);
            fm = fm * clampf(fpm_expf(ex), 0.0f, 1.0f);
        }
    }
    if (a.wet_mode != FP_WET_OFF) {
        const float nxm = a.nxm, nym = a.nym;
        const float hx = (float)s.i(F_CELL_X)[j] + s.f(F_POS_X)[j], hy = (float)s.i(F_CELL_Y)[j] + s.f(F_POS_Y)[j];
        const uint32_t pidx = (uint32_t)floorf(clampf(hx, 0.0f, nxm)) * a.ny + (uint32_t)floorf(clampf(hy, 0.0f, nym));
        float lam, frac;
        if (a.wet_mode == FP_WET_PRECIP_FIELD) {
            const float pr = pos_or_zero(a.precip[pidx]);
            lam = 0.0f;
            if (pr > 0.0f) lam = pos_or_zero(pos_or_zero(a.wet_a) * fpm_powf(pr, pos_or_zero(a.wet_b)));
            frac = a.wet_fraction;
        } else {
            lam = a.wet_lambda;
            frac = a.wet_fraction;
        }
        const float l = fmaxf(lam, 0.0f), dt = fabsf(a.dt), fr = clampf(frac, 0.0f, 1.0f);
        float prob = 0.0f;
        if (!(l <= 0.0f || dt <= 0.0f || fr <= 0.0f)) prob = clampf((1.0f - fpm_expf(-l * dt)) * fr, 0.0f, 1.0f);
        fm = fm * (1.0f - prob);
    }
    return fm;
}

// Mass adjoint of one step (emission gradients only).  For every particle j
// alive after step k (post state, post-compaction index): the observation at
// time k + 1 adds dC/dm, then the step's survival factor maps the mass
// adjoint to its slot at the start of the step (map[j], or j without
// compaction).
__global__ void __launch_bounds__(256) mass_adj_kernel(StepArgs a, Soa post, uint32_t n_post, const uint32_t *map,
                                                       bool observe, ObsArgs o, const double *rec_lam,
                                                       const float *lam_post, float *lam_pre)
{
    const uint64_t j = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (j >= n_post) return;
    const uint32_t p = map ? map[j] : (uint32_t)j;
    const bool act = (post.u(F_FLAGS)[j] & FLAG_ACTIVE) != 0u;
    float lm = lam_post[j];
    if (act && observe) {
        const Cic c = cic_of(o, (float)post.i(F_CELL_X)[j] + post.f(F_POS_X)[j],
                             (float)post.i(F_CELL_Y)[j] + post.f(F_POS_Y)[j], post.f(F_POS_Z)[j]);
        float g = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int32_t rec = o.node_rec[cic_node(o, c, k)];
            if (rec >= 0) g += cic_weight(c, k) * (float)rec_lam[rec];
        }
        lm += g;
    }
    // a particle inactive after the step either was inactive before it
    // (untouched: the adjoint passes through) or was deactivated by it (held
    // fixed: it carries no adjoint to earlier steps)
    lam_pre[p] = act ? lm * survival_factor(a, post, j) : (map ? 0.0f : lm);
}

uint32_t grid_for(uint64_t n) { return (uint32_t)((n + 255u) / 256u); }

}  // namespace

hipError_t launch_adj_step(const StepArgs &a, float *lam, uint64_t stride, double *au, double *av, hipStream_t st)
{
    if (a.n == 0) return hipSuccess;
    AdjArgs g{a, lam, stride, au, av};
    if (a.n_substeps == 4u) adj_step_kernel<4><<<grid_for(a.n), 256, 0, st>>>(g);
    else adj_step_kernel<-1><<<grid_for(a.n), 256, 0, st>>>(g);
    return hipGetLastError();
}

}  // namespace fp

// ====================================================================== host

struct fp_adj {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    fp_adj_desc d{};
    bool compact = false, emission_only = false;
    std::vector<uint32_t> counts;
    std::vector<uint64_t> offs;
    uint64_t n = 0, npts = 0, nxy = 0;

    float *met_base = nullptr;     // n_met_levels x {u, v, w} x npts
    float4 *met_packed = nullptr;  // (n_met_levels - 1) brackets x npts x 2
    float *level_heights = nullptr;
    std::vector<float> lh_host;
    uint32_t has_heights = fp::kLevelIndex;
    bool met_ready = false;
    float zmax = 0.0f;
    double level_seconds = 0.0;

    float *basis_u = nullptr, *basis_v = nullptr, *pert_u = nullptr, *pert_v = nullptr;
    bool basis_ready = false;

    float4 *pbl = nullptr;
    float *precip = nullptr, *pbl_stage = nullptr;
    bool pbl_ready = false, precip_ready = false;

    uint32_t *src_of = nullptr, *d_counts = nullptr;
    uint64_t *d_offs = nullptr;
    float *d_q = nullptr, *d_r = nullptr, *d_c = nullptr;

    fp::Soa work, work2;               // forward states between checkpoints
    fp::Soa final_state;               // where the last forward's final state is
    std::vector<uint32_t> ckpt_steps;  // steps whose state is kept (0, K, 2K, ..., n_steps)
    std::vector<fp::Soa> ckpt;
    std::vector<fp::Soa> tape;         // K - 1 sets
    float *lam = nullptr, *lam2 = nullptr;

    // per step of the last forward: active count at its start, Philox counter,
    // survivor map of the compaction after it
    std::vector<uint32_t> step_n;
    std::vector<std::array<uint32_t, 4>> step_ctr;
    std::vector<uint64_t> map_off;
    uint32_t *maps = nullptr;
    uint32_t *block_counts = nullptr, *active_dev = nullptr, *active_host = nullptr;

    int32_t *node_rec = nullptr;
    uint64_t n_nodes = 0;
    uint32_t n_rec = 0;
    double *rec_val = nullptr, *rec_lam = nullptr;
    std::vector<double> obs_host, val_host;

    double *au = nullptr, *av = nullptr, *gq = nullptr, *gr = nullptr, *gc = nullptr;
    fp_particle *aos_stage = nullptr;

    fp_step_params params{};
    bool forward_done = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double fwd_ms = 0.0, bwd_ms = 0.0;
};

namespace {

thread_local std::string g_adj_error;

int afail(fp_adj *e, int code, const std::string &msg)
{
    if (e) e->err = msg; else g_adj_error = msg;
    return code;
}

#define A_HIP(e, expr)                                                                        \
    do {                                                                                      \
        hipError_t h_ = (expr);                                                               \
        if (h_ != hipSuccess)                                                                 \
            return afail((e), h_ == hipErrorOutOfMemory ? FP_E_NOMEM : FP_E_DEVICE,           \
                         std::string(#expr) + ": " + hipGetErrorString(h_));                  \
    } while (0)
#define A_CHECK(e, cond, msg)                                                                 \
    do {                                                                                      \
        if (!(cond)) return afail((e), FP_E_INVALID, (msg));                                  \
    } while (0)
#define A_STATE(e, cond, msg)                                                                 \
    do {                                                                                      \
        if (!(cond)) return afail((e), FP_E_STATE, (msg));                                    \
    } while (0)

template <typename T>
hipError_t amalloc(T **p, uint64_t count)
{
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

void destroy(fp_adj *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    auto fr = [](void *p) { if (p) (void)hipFree(p); };
    fr(e->met_base); fr(e->met_packed); fr(e->level_heights);
    fr(e->basis_u); fr(e->basis_v); fr(e->pert_u); fr(e->pert_v);
    fr(e->pbl); fr(e->precip); fr(e->pbl_stage);
    fr(e->src_of); fr(e->d_counts); fr(e->d_offs); fr(e->d_q); fr(e->d_r); fr(e->d_c);
    fr(e->work.base); fr(e->work2.base);
    for (auto &s : e->ckpt) fr(s.base);
    for (auto &s : e->tape) fr(s.base);
    fr(e->lam); fr(e->lam2); fr(e->maps); fr(e->block_counts); fr(e->active_dev);
    if (e->active_host) (void)hipHostFree(e->active_host);
    fr(e->node_rec); fr(e->rec_val); fr(e->rec_lam);
    fr(e->au); fr(e->av); fr(e->gq); fr(e->gr); fr(e->gc); fr(e->aos_stage);
    if (e->e0) (void)hipEventDestroy(e->e0);
    if (e->e1) (void)hipEventDestroy(e->e1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

// copy the fields a step reads and writes (one species) of the first cnt slots
hipError_t copy_state(fp_adj *e, const fp::Soa &dst, const fp::Soa &src, uint64_t cnt)
{
    if (cnt == 0) return hipSuccess;
    for (uint32_t f : fp::kStateFields) {
        const hipError_t h = hipMemcpyAsync(dst.u(f), src.u(f), cnt * 4u, hipMemcpyDeviceToDevice, e->stream);
        if (h != hipSuccess) return h;
    }
    return hipSuccess;
}

// the step arguments of step k on state s (count and counter of the last forward)
fp::StepArgs step_args(const fp_adj *e, uint32_t k, const fp::Soa &s)
{
    fp::StepArgs a{};
    a.s = s;
    a.n = e->step_n[k];
    a.n_species = 1;
    const double t = (double)k * (double)e->params.dt_seconds;
    const uint32_t nb = e->d.n_met_levels - 1u;
    uint32_t b = (uint32_t)std::min<double>(std::floor(t / e->level_seconds), (double)(nb - 1u));
    a.met = e->met_packed + (uint64_t)b * e->npts * 2u;
    a.nx = e->d.met_nx;
    a.ny = e->d.met_ny;
    a.nz = e->d.met_nz;
    a.level_heights = e->level_heights;
    a.has_heights = e->has_heights;
    a.zmax = e->zmax;
    a.pbl = e->pbl;
    a.precip = e->precip;
    fp_step_params p = e->params;
    // interpolation_alpha over the bracket (timeloop.rs:2086-2120), clamped
    const double t0 = (double)b * e->level_seconds;
    p.alpha = (float)((t - t0) / e->level_seconds);
    for (int c = 0; c < 4; ++c) p.counter[c] = e->step_ctr[k][c];
    p.grid_fused = 0;
    fp::fill_step_scalars(a, p);
    a.ppt = 1;  // one particle per thread: the 5-wavefront single-row kernels (10 % faster on the 50M run)
    return a;
}

fp::ObsArgs obs_args(const fp_adj *e)
{
    return fp::ObsArgs{e->node_rec, e->d.met_nx, e->d.met_ny, e->d.obs_nz, e->d.obs_dz_m};
}

int ckpt_index(const fp_adj *e, uint32_t step)
{
    for (size_t j = 0; j < e->ckpt_steps.size(); ++j)
        if (e->ckpt_steps[j] == step) return (int)j;
    return -1;
}

bool observed(const fp_adj *e, uint32_t t) { return t >= e->d.obs_first_step && e->n_rec > 0; }

}  // namespace

extern "C" {

const char *fp_adj_last_error(const fp_adj *e) { return e ? e->err.c_str() : g_adj_error.c_str(); }

int fp_adj_create(int device_ordinal, const fp_adj_desc *desc, fp_adj **out)
{
    if (!desc || !out || !desc->particles_per_source) return afail(nullptr, FP_E_INVALID, "null argument");
    *out = nullptr;
    const fp_adj_desc &d = *desc;
    if (d.n_sources == 0) return afail(nullptr, FP_E_INVALID, "n_sources must be >= 1");
    if (d.met_nx < 2 || d.met_ny < 2 || d.met_nz == 0 || d.met_nz > fp::kMaxLevels)
        return afail(nullptr, FP_E_INVALID, "met grid must be at least 2 x 2 with 1..256 levels");
    if (d.n_met_levels < 2) return afail(nullptr, FP_E_INVALID, "n_met_levels must be >= 2");
    if (d.obs_nz < 2 || !(d.obs_dz_m > 0.0f)) return afail(nullptr, FP_E_INVALID, "obs lattice needs >= 2 levels and dz > 0");
    if (d.n_steps == 0 || d.checkpoint_every == 0) return afail(nullptr, FP_E_INVALID, "n_steps and checkpoint_every must be >= 1");
    if (d.obs_first_step == 0 || d.obs_first_step > d.n_steps)
        return afail(nullptr, FP_E_INVALID, "obs_first_step must be in [1, n_steps]");
    if (d.flags & ~(FP_ADJ_COMPACT | FP_ADJ_EMISSION_ONLY)) return afail(nullptr, FP_E_INVALID, "unknown flags");
    uint64_t total = 0;
    for (uint32_t k = 0; k < d.n_sources; ++k) {
        if (d.particles_per_source[k] == 0) return afail(nullptr, FP_E_INVALID, "every source needs >= 1 particle");
        total += d.particles_per_source[k];
    }
    if (total > fp::kMaxSlots) return afail(nullptr, FP_E_INVALID, "total particle count must be <= 2^30");
    const uint64_t npts = (uint64_t)d.met_nx * d.met_ny * d.met_nz;
    if (npts > fp::kMaxMetPoints) return afail(nullptr, FP_E_INVALID, "met grid too large (at most 2^27 points)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return afail(nullptr, FP_E_DEVICE, "no HIP device");
    if (device_ordinal < 0 || device_ordinal >= ndev) return afail(nullptr, FP_E_INVALID, "device ordinal out of range");
    if (hipSetDevice(device_ordinal) != hipSuccess) return afail(nullptr, FP_E_DEVICE, "hipSetDevice failed");

    fp_adj *e = new (std::nothrow) fp_adj();
    if (!e) return afail(nullptr, FP_E_NOMEM, "host allocation failed");
    e->device = device_ordinal;
    e->d = d;
    e->d.particles_per_source = nullptr;
    e->compact = (d.flags & FP_ADJ_COMPACT) != 0;
    e->emission_only = (d.flags & FP_ADJ_EMISSION_ONLY) != 0;
    e->counts.assign(d.particles_per_source, d.particles_per_source + d.n_sources);
    e->offs.resize(d.n_sources);
    uint64_t acc = 0;
    for (uint32_t k = 0; k < d.n_sources; ++k) { e->offs[k] = acc; acc += e->counts[k]; }
    e->n = total;
    e->npts = npts;
    e->nxy = (uint64_t)d.met_nx * d.met_ny;
    for (uint32_t k = 0; k < d.n_steps; k += d.checkpoint_every) e->ckpt_steps.push_back(k);
    e->ckpt_steps.push_back(d.n_steps);
    e->n_nodes = e->nxy * d.obs_nz;
    e->step_n.assign(d.n_steps + 1u, 0u);
    e->step_ctr.assign(d.n_steps + 1u, std::array<uint32_t, 4>{0u, 0u, 0u, 0u});
    e->map_off.assign(d.n_steps + 1u, 0u);

    auto bail = [&](hipError_t h, const char *what) {
        g_adj_error = std::string(what) + ": " + hipGetErrorString(h);
        destroy(e);
        return h == hipErrorOutOfMemory ? FP_E_NOMEM : FP_E_DEVICE;
    };
#define C_HIP(expr)                                   \
    do {                                              \
        hipError_t h_ = (expr);                       \
        if (h_ != hipSuccess) return bail(h_, #expr); \
    } while (0)
    C_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    C_HIP(hipEventCreate(&e->e0));
    C_HIP(hipEventCreate(&e->e1));
    C_HIP(amalloc(&e->met_base, (uint64_t)d.n_met_levels * 3u * npts));
    C_HIP(amalloc(&e->met_packed, (uint64_t)(d.n_met_levels - 1u) * npts * 2u));
    C_HIP(amalloc(&e->level_heights, (uint64_t)fp::level_table_words(d.met_nz)));
    if (d.n_modes) {
        C_HIP(amalloc(&e->basis_u, (uint64_t)d.n_modes * e->nxy));
        C_HIP(amalloc(&e->basis_v, (uint64_t)d.n_modes * e->nxy));
        C_HIP(amalloc(&e->d_c, (uint64_t)d.n_modes));
        C_HIP(amalloc(&e->gc, (uint64_t)d.n_modes));
    }
    C_HIP(amalloc(&e->pert_u, e->nxy));
    C_HIP(amalloc(&e->pert_v, e->nxy));
    C_HIP(amalloc(&e->pbl, e->nxy * 2u));
    C_HIP(amalloc(&e->precip, e->nxy));
    C_HIP(amalloc(&e->pbl_stage, e->nxy * 4u));
    C_HIP(amalloc(&e->src_of, total));
    C_HIP(amalloc(&e->d_counts, (uint64_t)d.n_sources));
    C_HIP(amalloc(&e->d_offs, (uint64_t)d.n_sources));
    C_HIP(amalloc(&e->d_q, (uint64_t)d.n_sources));
    C_HIP(amalloc(&e->d_r, (uint64_t)d.n_sources * 3u));
    C_HIP(amalloc(&e->gq, (uint64_t)d.n_sources));
    C_HIP(amalloc(&e->gr, (uint64_t)d.n_sources * 3u));
    auto soa_alloc = [&](fp::Soa &s) {
        s.capacity = total;
        return amalloc(&s.base, total * fp::F_COUNT);
    };
    C_HIP(soa_alloc(e->work));
    if (e->compact) C_HIP(soa_alloc(e->work2));  // the compacting loop alternates two sets
    e->ckpt.resize(e->ckpt_steps.size());
    for (auto &s : e->ckpt) C_HIP(soa_alloc(s));
    e->tape.resize(d.checkpoint_every > 1 ? d.checkpoint_every - 1u : 0u);
    for (auto &s : e->tape) C_HIP(soa_alloc(s));
    C_HIP(amalloc(&e->lam, total * fp::kLam));
    C_HIP(amalloc(&e->lam2, total * fp::kLam));
    if (e->compact) {
        C_HIP(amalloc(&e->maps, total * (uint64_t)d.n_steps));
        C_HIP(amalloc(&e->block_counts, fp::compact_blocks((uint32_t)total) + 1u));
        C_HIP(amalloc(&e->active_dev, 1));
        C_HIP(hipHostMalloc(reinterpret_cast<void **>(&e->active_host), 4, hipHostMallocDefault));
    }
    C_HIP(amalloc(&e->node_rec, e->n_nodes));
    C_HIP(amalloc(&e->rec_val, e->n_nodes));
    C_HIP(amalloc(&e->rec_lam, e->n_nodes));
    C_HIP(amalloc(&e->au, e->nxy));
    C_HIP(amalloc(&e->av, e->nxy));
    C_HIP(amalloc(&e->aos_stage, total));
    // zero every particle set once so that fields the step never writes read as zero
    for (fp::Soa *s : {&e->work, &e->work2})
        if (s->base) C_HIP(hipMemsetAsync(s->base, 0, total * fp::F_COUNT * 4u, e->stream));
    for (auto &s : e->ckpt) C_HIP(hipMemsetAsync(s.base, 0, total * fp::F_COUNT * 4u, e->stream));
    for (auto &s : e->tape) C_HIP(hipMemsetAsync(s.base, 0, total * fp::F_COUNT * 4u, e->stream));
    C_HIP(hipMemsetAsync(e->node_rec, 0xff, e->n_nodes * 4u, e->stream));
    std::vector<uint32_t> src(total);
    for (uint32_t k = 0; k < d.n_sources; ++k)
        std::fill(src.begin() + e->offs[k], src.begin() + e->offs[k] + e->counts[k], k);
    C_HIP(hipMemcpyAsync(e->src_of, src.data(), total * 4u, hipMemcpyHostToDevice, e->stream));
    C_HIP(hipMemcpyAsync(e->d_counts, e->counts.data(), d.n_sources * 4u, hipMemcpyHostToDevice, e->stream));
    C_HIP(hipMemcpyAsync(e->d_offs, e->offs.data(), d.n_sources * 8u, hipMemcpyHostToDevice, e->stream));
    C_HIP(hipStreamSynchronize(e->stream));
#undef C_HIP
    *out = e;
    return FP_OK;
}

int fp_adj_destroy(fp_adj *e)
{
    destroy(e);
    return FP_OK;
}

int fp_adj_met_upload(fp_adj *e, const float *u, const float *v, const float *w, const float *lh,
                      double seconds_between_levels)
{
    A_CHECK(e, e && u && v && w && lh, "null argument");
    A_CHECK(e, seconds_between_levels > 0.0 && std::isfinite(seconds_between_levels),
            "seconds_between_levels must be > 0");
    const uint32_t nz = e->d.met_nz;
    for (uint32_t k = 0; k < nz; ++k) A_CHECK(e, std::isfinite(lh[k]), "level heights must be finite");
    const bool metres = lh[nz - 1] > 0.0f;
    if (metres)
        for (uint32_t k = 1; k < nz; ++k) A_CHECK(e, lh[k] >= lh[k - 1], "level heights must be non-decreasing");
    e->has_heights = fp::level_table(lh, nz, e->lh_host);
    e->zmax = metres ? lh[nz - 1] : (float)(nz - 1);
    e->level_seconds = seconds_between_levels;
    const uint64_t np = e->npts;
    for (uint32_t l = 0; l < e->d.n_met_levels; ++l) {
        const float *f[3] = {u + (uint64_t)l * np, v + (uint64_t)l * np, w + (uint64_t)l * np};
        for (int c = 0; c < 3; ++c)
            A_HIP(e, hipMemcpyAsync(e->met_base + ((uint64_t)l * 3u + c) * np, f[c], np * 4u, hipMemcpyHostToDevice,
                                    e->stream));
    }
    A_HIP(e, hipMemcpyAsync(e->level_heights, e->lh_host.data(), e->lh_host.size() * 4u, hipMemcpyHostToDevice,
                            e->stream));
    A_HIP(e, hipStreamSynchronize(e->stream));
    e->met_ready = true;
    e->forward_done = false;
    return FP_OK;
}

int fp_adj_basis_upload(fp_adj *e, const float *bu, const float *bv)
{
    A_CHECK(e, e && bu && bv, "null argument");
    A_CHECK(e, e->d.n_modes > 0, "the run has no wind coefficients");
    const uint64_t cnt = (uint64_t)e->d.n_modes * e->nxy;
    A_HIP(e, hipMemcpyAsync(e->basis_u, bu, cnt * 4u, hipMemcpyHostToDevice, e->stream));
    A_HIP(e, hipMemcpyAsync(e->basis_v, bv, cnt * 4u, hipMemcpyHostToDevice, e->stream));
    A_HIP(e, hipStreamSynchronize(e->stream));
    e->basis_ready = true;
    e->forward_done = false;
    return FP_OK;
}

int fp_adj_pbl_upload(fp_adj *e, const float *ustar, const float *wstar, const float *hmix, const float *oli,
                      const float *precip)
{
    A_CHECK(e, e && ustar && wstar && hmix && oli, "null argument");
    const uint64_t np = e->nxy;
    const float *f[4] = {ustar, wstar, hmix, oli};
    for (int k = 0; k < 4; ++k)
        A_HIP(e, hipMemcpyAsync(e->pbl_stage + k * np, f[k], np * 4u, hipMemcpyHostToDevice, e->stream));
    A_HIP(e, fp::launch_pack_pbl(e->pbl_stage, e->pbl, (uint32_t)np, e->stream));
    if (precip) {
        A_HIP(e, hipMemcpyAsync(e->precip, precip, np * 4u, hipMemcpyHostToDevice, e->stream));
        e->precip_ready = true;
    }
    A_HIP(e, hipStreamSynchronize(e->stream));
    e->pbl_ready = true;
    e->forward_done = false;
    return FP_OK;
}

int fp_adj_receptors(fp_adj *e, uint32_t n_rec, const uint32_t *nodes, const double *observed)
{
    A_CHECK(e, e && (nodes || n_rec == 0), "null argument");
    A_CHECK(e, n_rec <= e->n_nodes, "more receptors than lattice nodes");
    std::vector<int32_t> table(e->n_nodes, -1);
    for (uint32_t r = 0; r < n_rec; ++r) {
        const uint32_t i = nodes[3u * r], j = nodes[3u * r + 1u], l = nodes[3u * r + 2u];
        A_CHECK(e, i < e->d.met_nx && j < e->d.met_ny && l < e->d.obs_nz, "receptor node outside the lattice");
        const uint64_t node = ((uint64_t)i * e->d.met_ny + j) * e->d.obs_nz + l;
        A_CHECK(e, table[node] < 0, "two receptors on one lattice node");
        table[node] = (int32_t)r;
    }
    A_HIP(e, hipMemcpyAsync(e->node_rec, table.data(), e->n_nodes * 4u, hipMemcpyHostToDevice, e->stream));
    A_HIP(e, hipStreamSynchronize(e->stream));
    e->n_rec = n_rec;
    e->obs_host.assign(n_rec, 0.0);
    if (observed) std::copy(observed, observed + n_rec, e->obs_host.begin());
    e->forward_done = false;
    return FP_OK;
}

int fp_adj_forward(fp_adj *e, const fp_step_params *p, const float *q, const float *r, const float *c,
                   double *receptor_values, double *loss)
{
    A_CHECK(e, e && p && q && r && (c || e->d.n_modes == 0), "null argument");
    A_STATE(e, e->met_ready, "fp_adj_forward before fp_adj_met_upload");
    A_STATE(e, e->pbl_ready, "fp_adj_forward before fp_adj_pbl_upload");
    A_STATE(e, e->d.n_modes == 0 || e->basis_ready, "fp_adj_forward before fp_adj_basis_upload");
    A_CHECK(e, std::isfinite(p->dt_seconds) && p->dt_seconds > 0.0f, "invalid dt_seconds");
    A_CHECK(e, p->n_substeps <= 4, "n_substeps must be 0..4");
    A_CHECK(e, p->direction == 1 || p->direction == -1, "direction must be +1 or -1");
    A_CHECK(e, p->dry_mode <= FP_DRY_UNIFORM, "differentiable runs support dry_mode off or uniform");
    A_CHECK(e, p->wet_mode != FP_WET_PER_SLOT && p->wet_mode <= FP_WET_PRECIP_FIELD, "unsupported wet_mode");
    A_STATE(e, p->wet_mode != FP_WET_PRECIP_FIELD || e->precip_ready, "precip field missing");
    A_CHECK(e, p->loss_mode <= FP_LOSS_MIN_MASS, "unknown loss_mode");
    for (uint32_t k = 0; k < e->d.n_sources; ++k)
        A_CHECK(e, std::isfinite(q[k]) && std::isfinite(r[3 * k]) && std::isfinite(r[3 * k + 1]) &&
                       std::isfinite(r[3 * k + 2]),
                "emissions and positions must be finite");
    e->params = *p;
    const uint32_t S = e->d.n_sources, K = e->d.n_modes;
    hipStream_t st = e->stream;
    A_HIP(e, hipMemcpyAsync(e->d_q, q, S * 4u, hipMemcpyHostToDevice, st));
    A_HIP(e, hipMemcpyAsync(e->d_r, r, S * 12u, hipMemcpyHostToDevice, st));
    if (K) A_HIP(e, hipMemcpyAsync(e->d_c, c, K * 4u, hipMemcpyHostToDevice, st));
    A_HIP(e, hipEventRecord(e->e0, st));
    // met with the perturbation
    if (K) fp::perturb_kernel<<<fp::grid_for(e->nxy), 256, 0, st>>>(e->basis_u, e->basis_v, e->d_c, K,
                                                                     (uint32_t)e->nxy, e->pert_u, e->pert_v);
    for (uint32_t b = 0; b + 1u < e->d.n_met_levels; ++b)
        fp::pack_bracket_kernel<<<fp::grid_for(e->npts), 256, 0, st>>>(
            e->met_base, e->npts, e->d.met_nz, b, K ? e->pert_u : nullptr, K ? e->pert_v : nullptr,
            e->met_packed + (uint64_t)b * e->npts * 2u);
    A_HIP(e, hipGetLastError());
    // sources: the state at step 0 is the first checkpoint
    fp::init_sources_kernel<<<fp::grid_for(e->n), 256, 0, st>>>(e->ckpt[0], (uint32_t)e->n, e->src_of,
                                                               e->d_counts, e->d_q, e->d_r, e->d.met_nx,
                                                               e->d.met_ny);
    fp::fill_f64_kernel<<<fp::grid_for(std::max<uint64_t>(e->n_rec, 1)), 256, 0, st>>>(e->rec_val, e->n_rec, 0.0);
    A_HIP(e, hipGetLastError());
    const fp::ObsArgs oa = obs_args(e);
    e->step_n[0] = (uint32_t)e->n;
    for (int cc = 0; cc < 4; ++cc) e->step_ctr[0][cc] = p->counter[cc];
    // A step reads the state where the previous one left it and writes it
    // where it is needed next, so no state is copied: a step whose end is a
    // checkpoint writes into that checkpoint, the step after it reads from
    // there (launch_step_out_of_place), the others step in place.
    // With compaction after every step (timeloop.rs:1185-1197), the records
    // are not moved by the compaction itself: step k + 1 reads survivor j of
    // step k through the saved map (launch_step_gathered) and writes it to
    // slot j, which is the state and the Philox slot a moved record would have.
    // Only the last compaction moves the records, so that the final state is
    // the compacted one.  The out-of-place and gathered steps write the words
    // a step changes; every other word of a record is zero here in every set
    // (sources start with zero words, the sets are zeroed at creation and no
    // kernel writes those words), so the records are whole.
    fp::Soa cur = e->ckpt[0];
    uint64_t moff = 0;
    for (uint32_t k = 0; k < e->d.n_steps; ++k) {
        const uint32_t nk = e->step_n[k];
        const int j = ckpt_index(e, k + 1u);
        const bool last = k + 1u == e->d.n_steps;
        fp::Soa dst;
        if (e->compact) dst = cur.base == e->work.base ? e->work2 : e->work;
        else dst = j >= 0 ? e->ckpt[j] : e->work;
        fp::StepArgs a = step_args(e, k, dst);
        if (dst.base == cur.base) {
            A_HIP(e, fp::launch_step(a, 256, st));
        } else if (e->compact && k > 0) {
            a.gather_map = e->maps + e->map_off[k - 1u];
            a.src = cur;
            A_HIP(e, fp::launch_step_gathered(a, false, st));
        } else {
            a.src = cur;
            A_HIP(e, fp::launch_step_out_of_place(a, st));
        }
        cur = dst;
        // cur: the state after step k (before its compaction)
        if (observed(e, k + 1u) && nk)
            fp::obs_deposit_kernel<<<fp::grid_for(nk), 256, 0, st>>>(cur, nk, oa, e->rec_val);
        uint32_t next = nk;
        if (e->compact) {
            e->map_off[k] = moff;
            if (last) {
                const fp::Soa out = cur.base == e->work.base ? e->work2 : e->work;
                A_HIP(e, fp::launch_compact(cur, out, nk, e->block_counts, e->maps + moff, e->active_dev, st));
                cur = out;
            } else {
                A_HIP(e, fp::launch_survivor_map(cur, nk, e->block_counts, e->maps + moff, e->active_dev, st));
            }
            A_HIP(e, hipMemcpyAsync(e->active_host, e->active_dev, 4, hipMemcpyDeviceToHost, st));
            A_HIP(e, hipStreamSynchronize(st));
            next = *e->active_host;
            moff += next;
        }
        e->step_n[k + 1u] = next;
        fp::counter_advance(e->step_ctr[k].data(), fp::counter_per_step(nk, p->n_substeps),
                            e->step_ctr[k + 1u].data());
        if (j >= 0 && e->compact) {
            if (!last) {
                if (next)
                    fp::gather_state_kernel<<<fp::grid_for(next), 256, 0, st>>>(cur, e->ckpt[j],
                                                                              e->maps + e->map_off[k], next);
            } else {
                A_HIP(e, copy_state(e, e->ckpt[j], cur, next));
            }
        }
    }
    e->final_state = cur;
    A_HIP(e, hipGetLastError());
    A_HIP(e, hipEventRecord(e->e1, st));
    e->val_host.assign(e->n_rec, 0.0);
    if (e->n_rec)
        A_HIP(e, hipMemcpyAsync(e->val_host.data(), e->rec_val, e->n_rec * 8u, hipMemcpyDeviceToHost, st));
    A_HIP(e, hipStreamSynchronize(st));
    float ms = 0.0f;
    A_HIP(e, hipEventElapsedTime(&ms, e->e0, e->e1));
    e->fwd_ms = ms;
    double L = 0.0;
    for (uint32_t k = 0; k < e->n_rec; ++k) {
        const double dlt = e->val_host[k] - e->obs_host[k];
        L += dlt * dlt;
    }
    if (receptor_values) std::copy(e->val_host.begin(), e->val_host.end(), receptor_values);
    if (loss) *loss = L;
    e->forward_done = true;
    return FP_OK;
}

int fp_adj_backward(fp_adj *e, double *grad_q, double *grad_r, double *grad_c)
{
    A_CHECK(e, e, "null argument");
    A_STATE(e, e->forward_done, "fp_adj_backward before fp_adj_forward");
    hipStream_t st = e->stream;
    const uint32_t S = e->d.n_sources, K = e->d.n_modes;
    const bool full = !e->emission_only;
    const uint32_t nf = full ? fp::kLam : 1u;   // adjoint quantities carried
    // the mass adjoint lives in row L_M of the adjoint buffers
    std::vector<double> lamr(e->n_rec);
    for (uint32_t k = 0; k < e->n_rec; ++k) lamr[k] = 2.0 * (e->val_host[k] - e->obs_host[k]);
    if (e->n_rec)
        A_HIP(e, hipMemcpyAsync(e->rec_lam, lamr.data(), e->n_rec * 8u, hipMemcpyHostToDevice, st));
    A_HIP(e, hipEventRecord(e->e0, st));
    fp::fill_f32_kernel<<<fp::grid_for(e->n * fp::kLam), 256, 0, st>>>(e->lam, e->n * fp::kLam, 0.0f);
    if (full && K) {
        fp::fill_f64_kernel<<<fp::grid_for(e->nxy), 256, 0, st>>>(e->au, e->nxy, 0.0);
        fp::fill_f64_kernel<<<fp::grid_for(e->nxy), 256, 0, st>>>(e->av, e->nxy, 0.0);
    }
    A_HIP(e, hipGetLastError());
    const fp::ObsArgs oa = obs_args(e);
    float *lam = e->lam, *lam2 = e->lam2;
    const uint64_t stride = e->n;
    const uint32_t nseg = (uint32_t)e->ckpt_steps.size() - 1u;
    for (uint32_t seg = nseg; seg-- > 0;) {
        const uint32_t k0 = e->ckpt_steps[seg], k1 = e->ckpt_steps[seg + 1u];
        // recompute the segment, keeping the state at the start of each step
        // (after the previous step's compaction, replayed from its saved map)
        for (uint32_t k = k0 + 1u; k < k1; ++k) {
            const fp::Soa &prev = (k - 1u == k0) ? e->ckpt[seg] : e->tape[k - 2u - k0];
            const fp::Soa &cur = e->tape[k - 1u - k0];
            if (e->compact) {
                // only the survivors of step k - 1, each stepped at its own slot
                fp::StepArgs a = step_args(e, k - 1u, cur);
                a.n = e->step_n[k];
                a.gather_map = e->maps + e->map_off[k - 1u];
                a.src = prev;
                A_HIP(e, fp::launch_step_gathered(a, true, st));
            } else {
                fp::StepArgs a = step_args(e, k - 1u, cur);
                a.src = prev;
                A_HIP(e, fp::launch_step_out_of_place(a, st));
            }
        }
        for (uint32_t k = k1; k-- > k0;) {
            const fp::Soa &after = (k + 1u == k1) ? e->ckpt[seg + 1u] : e->tape[k - k0];
            const fp::Soa &before = (k == k0) ? e->ckpt[seg] : e->tape[k - 1u - k0];
            const uint32_t n_after = e->step_n[k + 1u], n_before = e->step_n[k];
            const uint32_t *map = e->compact ? e->maps + e->map_off[k] : nullptr;
            if (!full) {
                // observation at k + 1, survivor map and survival factor in one pass
                if (map) fp::fill_f32_kernel<<<fp::grid_for(n_before), 256, 0, st>>>(lam2 + fp::L_M * stride, n_before, 0.0f);
                const fp::StepArgs a = step_args(e, k, after);
                if (n_after)
                    fp::mass_adj_kernel<<<fp::grid_for(n_after), 256, 0, st>>>(
                        a, after, n_after, map, observed(e, k + 1u), oa, e->rec_lam, lam + fp::L_M * stride,
                        (map ? lam2 : lam) + fp::L_M * stride);
                if (map) std::swap(lam, lam2);
                continue;
            }
            // observation at time k + 1 (the state after step k; compaction
            // keeps every observed particle, so the post-compaction state serves)
            if (observed(e, k + 1u) && n_after)
                fp::obs_adjoint_kernel<<<fp::grid_for(n_after), 256, 0, st>>>(after, n_after, oa, e->rec_lam, lam,
                                                                              stride);
            if (map) {
                if (n_before)
                    fp::zero_rows_kernel<<<fp::grid_for((uint64_t)n_before * nf), 256, 0, st>>>(lam2, stride, nf,
                                                                                                n_before);
                if (n_after)
                    fp::uncompact_kernel<<<fp::grid_for(n_after), 256, 0, st>>>(lam, lam2, map, n_after, stride, nf);
                std::swap(lam, lam2);
            }
            A_HIP(e, fp::launch_adj_step(step_args(e, k, before), lam, stride, K ? e->au : nullptr,
                                         K ? e->av : nullptr, st));
        }
    }
    fp::source_grad_kernel<<<S, 256, 0, st>>>(lam, stride, e->d_offs, e->d_counts, e->d_r, e->d.met_nx,
                                              e->d.met_ny, e->gq, e->gr);
    if (full && K)
        fp::coeff_grad_kernel<<<K, 256, 0, st>>>(e->au, e->av, e->basis_u, e->basis_v, (uint32_t)e->nxy, e->gc);
    A_HIP(e, hipGetLastError());
    A_HIP(e, hipEventRecord(e->e1, st));
    if (grad_q) A_HIP(e, hipMemcpyAsync(grad_q, e->gq, S * 8u, hipMemcpyDeviceToHost, st));
    if (grad_r) A_HIP(e, hipMemcpyAsync(grad_r, e->gr, S * 24u, hipMemcpyDeviceToHost, st));
    if (grad_c && K) A_HIP(e, hipMemcpyAsync(grad_c, e->gc, K * 8u, hipMemcpyDeviceToHost, st));
    A_HIP(e, hipStreamSynchronize(st));
    if (!full) {
        // position and wind gradients are not computed in this mode
        if (grad_r) std::fill(grad_r, grad_r + 3u * S, 0.0);
        if (grad_c) std::fill(grad_c, grad_c + K, 0.0);
    }
    float ms = 0.0f;
    A_HIP(e, hipEventElapsedTime(&ms, e->e0, e->e1));
    e->bwd_ms = ms;
    return FP_OK;
}

int fp_adj_download_state(fp_adj *e, fp_particle *out, uint64_t n)
{
    A_CHECK(e, e && (out || n == 0), "null argument");
    A_CHECK(e, n <= e->n, "more records than particles");
    A_STATE(e, e->forward_done, "fp_adj_download_state before fp_adj_forward");
    if (n == 0) return FP_OK;
    A_HIP(e, fp::launch_soa_to_aos(e->final_state, e->aos_stage, n, e->stream));  // the backward leaves it
    A_HIP(e, hipMemcpyAsync(out, e->aos_stage, n * sizeof(fp_particle), hipMemcpyDeviceToHost, e->stream));
    A_HIP(e, hipStreamSynchronize(e->stream));
    return FP_OK;
}

int fp_adj_active_counts(fp_adj *e, uint64_t *counts)
{
    A_CHECK(e, e && counts, "null argument");
    A_STATE(e, e->forward_done, "fp_adj_active_counts before fp_adj_forward");
    for (uint32_t k = 0; k <= e->d.n_steps; ++k) counts[k] = e->step_n[k];
    return FP_OK;
}

int fp_adj_last_ms(fp_adj *e, double *forward_ms, double *backward_ms)
{
    A_CHECK(e, e, "null argument");
    if (forward_ms) *forward_ms = e->fwd_ms;
    if (backward_ms) *backward_ms = e->bwd_ms;
    return FP_OK;
}

}  // extern "C"
