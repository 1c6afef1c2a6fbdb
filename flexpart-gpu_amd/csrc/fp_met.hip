// fp_met.hip - met ingest on the device: the hybrid sigma-pressure to height
// transform of io/vertical_transform.rs.
//
// transform_hybrid_sigma_to_height (vertical_transform.rs:144-212) walks every
// column on the CPU: p = A + B ps per level, virtual temperature
// T (1 + 0.608 max(q, 0)), hypsometric layer thicknesses (the
// |Tv - Tv_old| > 0.2 switch of verttransform_ecmwf.f90) accumulated into
// geometric heights, then the horizontal mean height profile that becomes the
// grid's level heights.  Here one thread per column runs the level loop (the
// column's levels are contiguous, so a wave's loads walk 64 neighbouring
// columns), and the mean profile is the reference's sequential f32 sum in
// (i, j) order with one lane per level: lane k adds height[c*nz + k], so the
// lanes of a wave read one column's levels in one coalesced access per c.
// Errors keep the reference's precedence: any pressure error (first in
// (i, j, k) order) before any error of the height loop (first column in
// (i, j) order, its first failing level).
#include "fp_physics.hpp"

namespace fp {
namespace {

constexpr uint32_t kVtBlock = 256;
constexpr float kRAirVt = 287.05f;  // lib.rs:24
constexpr float kGaVt = 9.81f;      // lib.rs:26
constexpr float kHypsScale = kRAirVt / kGaVt;
constexpr float kTvSwitch = 0.2f;
constexpr float kTvFactor = 0.608f;
constexpr float kF32Eps = 1.1920929e-7f;  // f32::EPSILON

struct VtArgs {
    const float *ps, *t, *q, *ak, *bk;
    float *pressure, *height;
    unsigned long long *err_pressure, *err_height;  // atomic-min keys, ~0 = none
    uint32_t ncol, nz;
};

// key = column << 32 | level << 8 | code: the smallest key is the first error
// in the reference's loop order
__device__ __forceinline__ unsigned long long vt_key(uint32_t c, uint32_t k, uint32_t code)
{
    return ((unsigned long long)c << 32) | ((unsigned long long)k << 8) | code;
}

__device__ __forceinline__ bool vt_finite(float v) { return v == v && fabsf(v) <= F32_MAX; }

// virtual_temperature_k (vertical_transform.rs:272-292); false = invalid state
__device__ __forceinline__ bool virtual_temperature(float t, float q, float &tv)
{
    if (!vt_finite(t) || t <= 0.0f || !vt_finite(q)) return false;
    tv = t * (1.0f + kTvFactor * fmaxf(q, 0.0f));
    return true;
}

// hypsometric_layer_thickness_m (vertical_transform.rs:294-313)
__device__ __forceinline__ float layer_thickness(float p_prev, float p_cur, float tv_prev, float tv_cur)
{
    const float pr_ln = fpm_logf(p_prev / p_cur);
    const float dtv = tv_cur - tv_prev;
    const float tvr_ln = fpm_logf(tv_cur / tv_prev);
    if (fabsf(dtv) > kTvSwitch && fabsf(tvr_ln) > kF32Eps) return kHypsScale * pr_ln * dtv / tvr_ln;
    return kHypsScale * pr_ln * tv_cur;
}

__global__ void __launch_bounds__(kVtBlock) vt_column_kernel(VtArgs a)
{
    __shared__ float sak[kMaxLevels], sbk[kMaxLevels];
    for (uint32_t k = threadIdx.x; k < a.nz; k += kVtBlock) {
        sak[k] = a.ak[k];
        sbk[k] = a.bk[k];
    }
    __syncthreads();
    const uint32_t c = blockIdx.x * kVtBlock + threadIdx.x;
    if (c >= a.ncol) return;
    const uint64_t base = (uint64_t)c * a.nz;
    const float ps = a.ps[c];
    // compute_hybrid_level_pressure_pa (vertical_transform.rs:106-137)
    bool perr = false;
    for (uint32_t k = 0; k < a.nz; ++k) {
        const float p = sak[k] + sbk[k] * ps;
        if (a.pressure) a.pressure[base + k] = p;
        if (!perr && (!vt_finite(p) || p <= 0.0f)) {
            atomicMin(a.err_pressure, vt_key(c, k, 1u));
            perr = true;
        }
    }
    if (perr) return;
    // the height loop (vertical_transform.rs:160-201)
    float p_prev = sak[0] + sbk[0] * ps, tv_prev, z = 0.0f;
    a.height[base] = 0.0f;
    if (!virtual_temperature(a.t[base], a.q ? a.q[base] : 0.0f, tv_prev)) {
        atomicMin(a.err_height, vt_key(c, 0u, 3u));
        return;
    }
    for (uint32_t k = 1; k < a.nz; ++k) {
        const float p = sak[k] + sbk[k] * ps;
        if (p >= p_prev) {
            atomicMin(a.err_height, vt_key(c, k, 2u));
            return;
        }
        float tv;
        if (!virtual_temperature(a.t[base + k], a.q ? a.q[base + k] : 0.0f, tv)) {
            atomicMin(a.err_height, vt_key(c, k, 3u));
            return;
        }
        const float zn = z + layer_thickness(p_prev, p, tv_prev, tv);
        if (!vt_finite(zn) || zn <= z) {
            atomicMin(a.err_height, vt_key(c, k, 4u));
            return;
        }
        a.height[base + k] = zn;
        z = zn;
        p_prev = p;
        tv_prev = tv;
    }
}

// horizontal_mean_height_profile_m (vertical_transform.rs:315-334): lane k of
// the launch sums level k over the columns in (i, j) order in f32, then
// divides by (f32)(nx*ny).  Four partial loads are in flight per step; the
// additions stay in order.
__global__ void __launch_bounds__(64) vt_mean_kernel(const float *height, uint32_t ncol, uint32_t nz, float *mean)
{
    const uint32_t k = blockIdx.x * 64u + threadIdx.x;
    if (k >= nz) return;
    float s = 0.0f;
    uint32_t c = 0;
    for (; c + 4u <= ncol; c += 4u) {
        const float h0 = height[(uint64_t)c * nz + k], h1 = height[(uint64_t)(c + 1u) * nz + k];
        const float h2 = height[(uint64_t)(c + 2u) * nz + k], h3 = height[(uint64_t)(c + 3u) * nz + k];
        s += h0;
        s += h1;
        s += h2;
        s += h3;
    }
    for (; c < ncol; ++c) s += height[(uint64_t)c * nz + k];
    mean[k] = s / (float)ncol;
}

}  // namespace

hipError_t launch_vertical_transform(const float *ps, const float *t, const float *q, const float *ak,
                                     const float *bk, float *pressure, float *height, float *mean,
                                     unsigned long long *err, uint32_t ncol, uint32_t nz, hipStream_t st)
{
    VtArgs a{ps, t, q, ak, bk, pressure, height, err, err + 1, ncol, nz};
    vt_column_kernel<<<(ncol + kVtBlock - 1) / kVtBlock, kVtBlock, 0, st>>>(a);
    if (hipError_t e = hipGetLastError()) return e;
    vt_mean_kernel<<<(nz + 63u) / 64u, 64, 0, st>>>(height, ncol, nz, mean);
    return hipGetLastError();
}

}  // namespace fp
