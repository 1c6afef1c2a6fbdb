// fp_sort.hip - Morton (Z-order) re-sort of the particle slots on the device.
//
// ParticleStore::sort_spatially (particles/mod.rs:649-699) on the host
// becomes: an f64 min/max reduction of the active positions (the
// ActiveParticleExtent bounds), a key per slot (quantise each axis to
// bits_per_axis bins, interleave x,y,z; inactive slots get a sentinel above
// every active key), a stable LSD radix sort of (key, slot) pairs over only
// the 3*bits+1 key bits, and a gather of every SoA field through the sorted
// slot list.  Stability plus the sentinel reproduce the reference order
// (active by key, ties and inactive slots in original order).
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>

#include "fp_internal.hpp"

namespace fp {
namespace {

constexpr uint32_t FLAG_ACTIVE = 1u;
constexpr uint32_t kRB = 256;

__device__ __forceinline__ void slot_pos(const Soa &s, uint64_t i, double &x, double &y, double &z)
{
    x = (double)s.i(F_CELL_X)[i] + (double)s.f(F_POS_X)[i];
    y = (double)s.i(F_CELL_Y)[i] + (double)s.f(F_POS_Y)[i];
    z = (double)s.f(F_POS_Z)[i];
}

// Per-block min/max of active positions: out[b*7 + {0..5}] = xmin,xmax,ymin,
// ymax,zmin,zmax; out[b*7+6] = 1 if the block saw an active slot.
__global__ void __launch_bounds__(kRB) bounds_partial_kernel(Soa s, uint32_t n, double *out)
{
    __shared__ double red[6][kRB];
    __shared__ uint32_t any;
    if (threadIdx.x == 0) any = 0;
    double v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    bool seen = false;
    for (uint64_t i = (uint64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kRB) {
        if (!(s.u(F_FLAGS)[i] & FLAG_ACTIVE)) continue;
        double x, y, z;
        slot_pos(s, i, x, y, z);
        v[0] = fmin(v[0], x); v[1] = fmax(v[1], x);
        v[2] = fmin(v[2], y); v[3] = fmax(v[3], y);
        v[4] = fmin(v[4], z); v[5] = fmax(v[5], z);
        seen = true;
    }
    __syncthreads();
    if (seen) atomicOr(&any, 1u);
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (uint32_t off = kRB / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            for (int k = 0; k < 6; k += 2) {
                red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + off]);
                red[k + 1][threadIdx.x] = fmax(red[k + 1][threadIdx.x], red[k + 1][threadIdx.x + off]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 6; ++k) out[blockIdx.x * 7 + k] = red[k][0];
        out[blockIdx.x * 7 + 6] = any ? 1.0 : 0.0;
    }
}

// One block folds the per-block partials (min/max are exact, so the order
// of reduction does not matter).
__global__ void __launch_bounds__(kRB) bounds_final_kernel(const double *part, uint32_t nparts,
                                                           double *bounds)
{
    __shared__ double red[7][kRB];
    double v[7] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY, 0.0};
    for (uint32_t b = threadIdx.x; b < nparts; b += kRB) {
        const double *p = part + b * 7;
        if (p[6] == 0.0) continue;
        v[6] = 1.0;
        for (int k = 0; k < 6; k += 2) {
            v[k] = fmin(v[k], p[k]);
            v[k + 1] = fmax(v[k + 1], p[k + 1]);
        }
    }
    for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (uint32_t off = kRB / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            for (int k = 0; k < 6; k += 2) {
                red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + off]);
                red[k + 1][threadIdx.x] = fmax(red[k + 1][threadIdx.x], red[k + 1][threadIdx.x + off]);
            }
            red[6][threadIdx.x] = fmax(red[6][threadIdx.x], red[6][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) bounds[threadIdx.x] = red[threadIdx.x][0];
}

// quantize_axis (particles/mod.rs:415-425)
__device__ __forceinline__ uint32_t quantize(double v, double lo, double hi, uint32_t max_bin)
{
    if (!isfinite(v) || !isfinite(lo) || !isfinite(hi)) return 0u;
    const double span = hi - lo;
    if (span <= 2.220446049250313e-16) return 0u;
    double t = (v - lo) / span;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return (uint32_t)floor(t * (double)max_bin);
}

// Spread the low 21 bits of v so that bit b moves to bit 3b.
__device__ __forceinline__ uint64_t spread3(uint32_t v)
{
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x001F00000000FFFFull;
    x = (x | (x << 16)) & 0x001F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

template <typename K>
__global__ void keys_kernel(Soa s, uint32_t n, const double *bounds, uint32_t bits, K *keys,
                            uint32_t *vals)
{
    const uint64_t i = (uint64_t)blockIdx.x * kRB + threadIdx.x;
    if (i >= n) return;
    const K sentinel = (K)1 << (3u * bits);
    K key = sentinel;
    if ((s.u(F_FLAGS)[i] & FLAG_ACTIVE) && bounds[6] != 0.0) {
        double x, y, z;
        slot_pos(s, i, x, y, z);
        const uint32_t mb = (1u << bits) - 1u;
        const uint64_t m = spread3(quantize(x, bounds[0], bounds[1], mb)) |
                           (spread3(quantize(y, bounds[2], bounds[3], mb)) << 1) |
                           (spread3(quantize(z, bounds[4], bounds[5], mb)) << 2);
        key = (K)m;
    }
    keys[i] = key;
    vals[i] = (uint32_t)i;
}

// Record gather: the permutation reads whole 96-byte particle records from an
// AoS copy (six 16-byte loads per slot, no partial cache lines wasted on
// random 4-byte reads) and writes each field of the sorted SoA coalesced.
__global__ void gather_records_kernel(const uint4 *aos, Soa out, uint32_t n, const uint32_t *new_to_old)
{
    const uint64_t j = (uint64_t)blockIdx.x * kRB + threadIdx.x;
    if (j >= n) return;
    const uint4 *rec = aos + (uint64_t)new_to_old[j] * (F_COUNT / 4);
    uint4 w[F_COUNT / 4];
#pragma unroll
    for (uint32_t q = 0; q < F_COUNT / 4; ++q) w[q] = rec[q];
#pragma unroll
    for (uint32_t q = 0; q < F_COUNT / 4; ++q) {
        out.u(4 * q + 0)[j] = w[q].x;
        out.u(4 * q + 1)[j] = w[q].y;
        out.u(4 * q + 2)[j] = w[q].z;
        out.u(4 * q + 3)[j] = w[q].w;
    }
}

template <typename K>
hipError_t sort_pairs(SortWork &w, uint32_t n, uint32_t end_bit, hipStream_t st)
{
    K *k_in = static_cast<K *>(w.keys);
    K *k_out = static_cast<K *>(w.keys_alt);
    size_t need = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, need, k_in, k_out, w.vals, w.vals_alt, (size_t)n, 0u,
                                             end_bit, st);
    if (e != hipSuccess) return e;
    if (need > w.temp_bytes) {
        if (w.temp) {
            e = hipStreamSynchronize(st);  // a previous sort may still use it
            if (e != hipSuccess) return e;
            (void)hipFree(w.temp);
        }
        w.temp = nullptr;
        w.temp_bytes = 0;
        e = hipMalloc(&w.temp, need);
        if (e != hipSuccess) return e;
        w.temp_bytes = need;
    }
    return rocprim::radix_sort_pairs(w.temp, need, k_in, k_out, w.vals, w.vals_alt, (size_t)n, 0u, end_bit, st);
}

}  // namespace

hipError_t sort_work_alloc(SortWork &w, uint64_t cap)
{
    cap = std::max<uint64_t>(cap, 1);
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&w.bounds_partial), kSortPartials * 7 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&w.bounds_dev), 8 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&w.keys, cap * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&w.keys_alt, cap * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&w.vals), cap * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&w.vals_alt), cap * sizeof(uint32_t));
    if (e == hipSuccess) w.cap = cap;
    return e;
}

void sort_work_free(SortWork &w)
{
    for (void *q : {(void *)w.bounds_partial, (void *)w.bounds_dev, w.keys, w.keys_alt, (void *)w.vals,
                    (void *)w.vals_alt, w.temp})
        if (q) (void)hipFree(q);
    w = SortWork{};
}

hipError_t sort_morton_range(SortWork &w, const fp_sort_params &p, const Soa &in, const Soa &out,
                             fp_particle *aos_stage, uint32_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    if (n > w.cap) return hipErrorInvalidValue;
    hipError_t e;
    if (p.explicit_bounds) {
        double b[7];
        for (int k = 0; k < 6; ++k) b[k] = p.bounds[k];
        b[6] = 1.0;
        e = hipMemcpyAsync(w.bounds_dev, b, sizeof b, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(st);  // b lives on this stack frame
        if (e != hipSuccess) return e;
    } else {
        const uint32_t nparts = (uint32_t)std::min<uint64_t>((n + kRB - 1) / kRB, kSortPartials);
        bounds_partial_kernel<<<nparts, kRB, 0, st>>>(in, n, w.bounds_partial);
        bounds_final_kernel<<<1, kRB, 0, st>>>(w.bounds_partial, nparts, w.bounds_dev);
    }
    const uint32_t grid = (n + kRB - 1) / kRB;
    const uint32_t end_bit = 3u * p.bits_per_axis + 1u;
    if (end_bit <= 32u) {
        keys_kernel<uint32_t><<<grid, kRB, 0, st>>>(in, n, w.bounds_dev, p.bits_per_axis,
                                                    static_cast<uint32_t *>(w.keys), w.vals);
        e = sort_pairs<uint32_t>(w, n, end_bit, st);
    } else {
        keys_kernel<uint64_t><<<grid, kRB, 0, st>>>(in, n, w.bounds_dev, p.bits_per_axis,
                                                    static_cast<uint64_t *>(w.keys), w.vals);
        e = sort_pairs<uint64_t>(w, n, end_bit, st);
    }
    if (e != hipSuccess) return e;
    e = launch_soa_to_aos(in, aos_stage, n, st);
    if (e != hipSuccess) return e;
    gather_records_kernel<<<grid, kRB, 0, st>>>(reinterpret_cast<const uint4 *>(aos_stage), out, n, w.vals_alt);
    return hipGetLastError();
}

}  // namespace fp
