// fp_met.hip - met ingest on the device: the hybrid sigma-pressure to height
// transform of io/vertical_transform.rs.
//
// transform_hybrid_sigma_to_height (vertical_transform.rs:144-212) walks every
// column on the CPU: p = A + B ps per level, virtual temperature
// T (1 + 0.608 max(q, 0)), hypsometric layer thicknesses (the
// |Tv - Tv_old| > 0.2 switch of verttransform_ecmwf.f90) accumulated into
// geometric heights, then the horizontal mean height profile that becomes the
// grid's level heights.  Here one thread per column runs the level loop, and
// the mean profile is the reference's sequential f32 sum in (i, j) order with
// one lane per level, fed from LDS tiles that other waves stream in.
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

// horizontal_mean_height_profile_m (vertical_transform.rs:315-334): the sum of
// level k over the columns in (i, j) order in f32, divided by (f32)(nx*ny).
// One block: lane k of the summing waves owns level k and adds the columns in
// order from an LDS tile (lanes read consecutive words of one column, no bank
// conflicts).  The other waves keep the next tile in registers: each iteration
// they store it to the free LDS buffer and issue the loads of the tile after,
// so a tile's HBM latency overlaps the summation of the previous one.
constexpr uint32_t kMeanBlock = 512;
constexpr uint32_t kMeanTileFloats = 8192;  // per LDS buffer (32 KB)
constexpr uint32_t kMeanPer = 32;           // tile floats per loader thread (>= 8192 / 256)

__global__ void __launch_bounds__(kMeanBlock) vt_mean_kernel(const float *height, uint32_t ncol, uint32_t nz,
                                                            float *mean)
{
    __shared__ float buf[2][kMeanTileFloats];
    const uint32_t tile = kMeanTileFloats / nz;  // columns per tile
    const uint32_t ntiles = (ncol + tile - 1) / tile;
    const uint32_t sum_threads = (nz + kWave - 1) / kWave * kWave;
    const uint32_t tid = threadIdx.x;
    const bool summer = tid < sum_threads;
    const uint32_t lid = tid - sum_threads, nload = kMeanBlock - sum_threads;
    float r[kMeanPer];
    auto tile_floats = [&](uint32_t tix) { return min(tile, ncol - tix * tile) * nz; };
    auto fetch = [&](uint32_t tix) {  // loaders: tile tix into registers
        const uint64_t base = (uint64_t)tix * tile * nz;
        const uint32_t n = tile_floats(tix);
#pragma unroll
        for (uint32_t q = 0; q < kMeanPer; ++q) {
            const uint32_t e = lid + q * nload;
            if (e < n) r[q] = height[base + e];
        }
    };
    auto put = [&](uint32_t tix, float *dst) {  // loaders: registers into LDS
        const uint32_t n = tile_floats(tix);
#pragma unroll
        for (uint32_t q = 0; q < kMeanPer; ++q) {
            const uint32_t e = lid + q * nload;
            if (e < n) dst[e] = r[q];
        }
    };
    {  // tile 0 by every thread
        const uint32_t n = tile_floats(0);
        for (uint32_t e = tid; e < n; e += kMeanBlock) buf[0][e] = height[e];
    }
    if (!summer && ntiles > 1) fetch(1);
    __syncthreads();
    float s = 0.0f;
    for (uint32_t tix = 0; tix < ntiles; ++tix) {
        if (!summer) {
            if (tix + 1 < ntiles) put(tix + 1, buf[(tix + 1) & 1]);
            if (tix + 2 < ntiles) fetch(tix + 2);
        } else if (tid < nz) {
            const float *b = buf[tix & 1] + tid;
            const uint32_t cnt = min(tile, ncol - tix * tile);
            uint32_t c = 0;
            for (; c + 16u <= cnt; c += 16u) {  // 16 LDS reads in flight, then the adds in order
                float h[16];
#pragma unroll
                for (uint32_t u = 0; u < 16u; ++u) h[u] = b[(c + u) * nz];
#pragma unroll
                for (uint32_t u = 0; u < 16u; ++u) s += h[u];
            }
            for (; c < cnt; ++c) s += b[c * nz];
        }
        __syncthreads();
    }
    if (tid < nz) mean[tid] = s / (float)ncol;
}

}  // namespace

hipError_t launch_vertical_transform(const float *ps, const float *t, const float *q, const float *ak,
                                     const float *bk, float *pressure, float *height, float *mean,
                                     unsigned long long *err, uint32_t ncol, uint32_t nz, hipStream_t st)
{
    VtArgs a{ps, t, q, ak, bk, pressure, height, err, err + 1, ncol, nz};
    vt_column_kernel<<<(ncol + kVtBlock - 1) / kVtBlock, kVtBlock, 0, st>>>(a);
    if (hipError_t e = hipGetLastError()) return e;
    vt_mean_kernel<<<1, kMeanBlock, 0, st>>>(height, ncol, nz, mean);
    return hipGetLastError();
}

}  // namespace fp
