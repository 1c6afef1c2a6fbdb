// fp_aux.hip - data movement kernels around the step: met/PBL packing,
// AoS <-> SoA transposes, stable stream compaction and the Philox
// known-answer kernel.
#include "fp_physics.hpp"

namespace fp {
namespace {

// 32-byte met record per grid point: {u0, v0, w0, u1, v1, w1, 0, 0}.  One
// record holds both time levels of all three components, so a trilinear
// corner is two 16-byte loads from one cache line.
__global__ void pack_met_kernel(const float *u0, const float *v0, const float *w0,
                                const float *u1, const float *v1, const float *w1, float4 *met,
                                uint32_t npts)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    met[2u * i] = make_float4(u0[i], v0[i], w0[i], u1[i]);
    met[2u * i + 1u] = make_float4(v1[i], w1[i], 0.0f, 0.0f);
}

// Two 16-byte records per PBL cell: {u*, w*, hmix, 1/L} as uploaded and the
// cell's height-independent Hanna terms (pbl_cell_terms_dev).
__global__ void pack_pbl_kernel(const float *stage, float4 *pbl, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pb = make_float4(stage[i], stage[n + i], stage[2u * n + i], stage[3u * n + i]);
    pbl[2u * i] = pb;
    pbl[2u * i + 1u] = pbl_cell_terms_dev(pb);
}

// AoS (96 B records) -> SoA.  slots == nullptr writes record r to slot r.
__global__ void aos_to_soa_kernel(const uint32_t *aos, const uint32_t *slots, uint64_t n, Soa s)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint64_t slot = slots ? slots[r] : r;
    const uint32_t *src = aos + r * F_COUNT;
#pragma unroll
    for (uint32_t f = 0; f < F_COUNT; ++f) s.u(f)[slot] = src[f];
}

// Tiled transposes through LDS for whole ranges: a block moves kT records;
// SoA side reads/writes each field coalesced, AoS side reads/writes the
// block's contiguous kT*96-byte span coalesced.
constexpr uint32_t kT = 256;

__global__ void __launch_bounds__(kT) soa_to_aos_tiled(Soa s, uint32_t *aos, uint64_t n)
{
    __shared__ uint32_t tile[F_COUNT][kT + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kT;
    const uint32_t t = threadIdx.x;
    const uint32_t cnt = (uint32_t)min<uint64_t>(kT, n - base);
    if (t < cnt)
#pragma unroll
        for (uint32_t f = 0; f < F_COUNT; ++f) tile[f][t] = s.u(f)[base + t];
    __syncthreads();
    uint32_t *dst = aos + base * F_COUNT;
    for (uint32_t w = t; w < cnt * F_COUNT; w += kT) dst[w] = tile[w % F_COUNT][w / F_COUNT];
}

__global__ void __launch_bounds__(kT) aos_to_soa_tiled(const uint32_t *aos, uint64_t n, Soa s)
{
    __shared__ uint32_t tile[F_COUNT][kT + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kT;
    const uint32_t t = threadIdx.x;
    const uint32_t cnt = (uint32_t)min<uint64_t>(kT, n - base);
    const uint32_t *src = aos + base * F_COUNT;
    for (uint32_t w = t; w < cnt * F_COUNT; w += kT) tile[w % F_COUNT][w / F_COUNT] = src[w];
    __syncthreads();
    if (t < cnt)
#pragma unroll
        for (uint32_t f = 0; f < F_COUNT; ++f) s.u(f)[base + t] = tile[f][t];
}

// ---------------------------------------------------------------- compaction
//
// Reduce-then-scan over the flag array (no inter-block waiting):
//   1. count_kernel  : active slots per tile of kTile slots
//   2. scan_kernel   : exclusive scan of the tile counts (one block), total
//   3. scatter_kernel: each tile ranks its active slots with wave ballots and
//                      copies them to offset + rank; slots at or beyond the
//                      total receive a deactivated copy of themselves
// The survivor order is ascending slot order, as in compaction.wgsl.
//
// The scatter is a 24-stream copy and HBM-bound: a block first ranks all of
// its kRows rows (one barrier), then copies field by field, issuing the loads
// of kFieldGroup fields x kRows rows before their stores, so each wave keeps
// 16 independent loads in flight instead of waiting on a barrier per row.

constexpr uint32_t kBlock = 256;
constexpr uint32_t kRows = 8;
constexpr uint32_t kTile = kBlock * kRows;
constexpr uint32_t kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kFieldGroup = 2;
static_assert(F_COUNT % kFieldGroup == 0, "field groups");

__device__ __forceinline__ uint32_t active_at(const Soa &s, uint64_t i, uint32_t n)
{
    return (i < n) ? (s.u(F_FLAGS)[i] & FLAG_ACTIVE) : 0u;
}

__global__ void __launch_bounds__(kBlock) count_kernel(Soa s, uint32_t n, uint32_t *tile_counts)
{
    __shared__ uint32_t wsum[kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    uint32_t c = 0;
#pragma unroll
    for (uint32_t r = 0; r < kRows; ++r) c += active_at(s, base + r * kBlock + threadIdx.x, n);
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
    if (__lane_id() == 0) wsum[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kWavesPerBlock; ++w) t += wsum[w];
        tile_counts[blockIdx.x] = t;
    }
}

// Exclusive scan of ntiles counts in place; writes the total.
__global__ void __launch_bounds__(1024) scan_kernel(uint32_t *tile_counts, uint32_t ntiles,
                                                    uint32_t *total)
{
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (ntiles + 1023u) / 1024u;
    const uint32_t lo = t * per, hi = min(lo + per, ntiles);
    uint32_t sum = 0;
    for (uint32_t k = lo; k < hi; ++k) sum += tile_counts[k];
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t k = lo; k < hi; ++k) {
        const uint32_t c = tile_counts[k];
        tile_counts[k] = run;
        run += c;
    }
    if (t == 1023u) *total = part[1023];
}

// COPY = false writes the survivor map only (the records stay where they are).
template <bool COPY>
__global__ void __launch_bounds__(kBlock) scatter_kernel(Soa in, Soa out, uint32_t n,
                                                         const uint32_t *tile_offsets,
                                                         const uint32_t *total_p,
                                                         uint32_t *survivor_map)
{
    __shared__ uint32_t wcount[kRows][kWavesPerBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kTile + threadIdx.x;
    const uint32_t wave = threadIdx.x / kWave, lane = __lane_id();
    const uint64_t below_mask = (1ull << lane) - 1ull;
    // ranks within each row's wave, then the row offsets of the waves
    uint32_t act = 0, dst[kRows];
#pragma unroll
    for (uint32_t r = 0; r < kRows; ++r) act |= active_at(in, base + r * kBlock, n) << r;
#pragma unroll
    for (uint32_t r = 0; r < kRows; ++r) {
        const uint64_t bal = __ballot((act >> r) & 1u);
        dst[r] = (uint32_t)__popcll(bal & below_mask);
        if (lane == 0) wcount[r][wave] = (uint32_t)__popcll(bal);
    }
    __syncthreads();
    const uint32_t total = *total_p;
    uint32_t running = tile_offsets[blockIdx.x];
#pragma unroll
    for (uint32_t r = 0; r < kRows; ++r) {
        uint32_t woff = 0, rowtot = 0;
#pragma unroll
        for (uint32_t w = 0; w < kWavesPerBlock; ++w) {
            const uint32_t c = wcount[r][w];
            woff += w < wave ? c : 0u;
            rowtot += c;
        }
        dst[r] += running + woff;
        running += rowtot;
    }
    // survivors to offset + rank; slots at or beyond the total (whether
    // active or not) also keep a deactivated copy of themselves
#pragma unroll
    for (uint32_t f0 = 0; f0 < (COPY ? F_COUNT : 0u); f0 += kFieldGroup) {
        uint32_t v[kFieldGroup][kRows];
#pragma unroll
        for (uint32_t g = 0; g < kFieldGroup; ++g)
#pragma unroll
            for (uint32_t r = 0; r < kRows; ++r) {
                const uint64_t i = base + r * kBlock;
                v[g][r] = i < n ? in.u(f0 + g)[i] : 0u;
            }
#pragma unroll
        for (uint32_t g = 0; g < kFieldGroup; ++g)
#pragma unroll
            for (uint32_t r = 0; r < kRows; ++r) {
                const uint64_t i = base + r * kBlock;
                const uint32_t f = f0 + g;
                if ((act >> r) & 1u) out.u(f)[dst[r]] = v[g][r];
                if (i < n && i >= total) out.u(f)[i] = f == F_FLAGS ? (v[g][r] & ~FLAG_ACTIVE) : v[g][r];
            }
    }
#pragma unroll
    for (uint32_t r = 0; r < kRows; ++r)
        if ((act >> r) & 1u) survivor_map[dst[r]] = (uint32_t)(base + r * kBlock);
}

__global__ void philox_kernel(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0_in,
                              uint32_t k1_in, uint64_t n, uint32_t *out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t lo = (uint64_t)c0 + (uint32_t)i;
    const uint64_t mid = (uint64_t)c1 + (uint32_t)(i >> 32) + (lo >> 32);
    const uint64_t hi = (uint64_t)c2 + (mid >> 32);
    uint32_t x = (uint32_t)lo, y = (uint32_t)mid, z = (uint32_t)hi, w = c3 + (uint32_t)(hi >> 32);
    uint32_t k0 = k0_in, k1 = k1_in;
    for (int r = 0; r < 10; ++r) {
        const uint32_t h0 = __umulhi(0xD2511F53u, x), l0 = 0xD2511F53u * x;
        const uint32_t h1 = __umulhi(0xCD9E8D57u, z), l1 = 0xCD9E8D57u * z;
        x = h1 ^ y ^ k0; y = l1; z = h0 ^ w ^ k1; w = l0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[4 * i + 0] = x; out[4 * i + 1] = y; out[4 * i + 2] = z; out[4 * i + 3] = w;
}

__global__ void math_probe_kernel(int fn, const float *x, const float *y, float *out, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = x[i], b = y[i];
    float s, c, r;
    switch (fn) {
    case FP_PROBE_EXP: r = fpm_expf(a); break;
    case FP_PROBE_LOG: r = fpm_logf(a); break;
    case FP_PROBE_SIN: fpm_sincosf(a, &s, &c); r = s; break;
    case FP_PROBE_COS: fpm_sincosf(a, &s, &c); r = c; break;
    case FP_PROBE_POW: r = fpm_powf(a, b); break;
    case FP_PROBE_SQRT: r = fpm_sqrtf(a); break;
    case FP_PROBE_DIV: r = a / b; break;
    case FP_PROBE_LOG_UNIT: r = fpm_logf_unit(a); break;
    default: r = 0.0f;
    }
    out[i] = r;
}

__global__ void __launch_bounds__(256) level_probe_kernel(const float *table, uint32_t nz, uint32_t mode,
                                                         const float *z, float *out, uint64_t n)
{
    __shared__ float lh[kLevelTableMax];
    for (uint32_t k = threadIdx.x; k < level_table_words(nz); k += 256) lh[k] = table[k];
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = level_of(z[i], lh, nz, mode);
}

// Exhaustive check of the step's square root (sqrt_in_range) over the float
// bit patterns [first, last]: the result s is the correctly rounded root of x
// exactly when x lies strictly between the squares of the midpoints from s to
// its two float neighbours.  Those midpoints have at most 26 significant bits,
// so their binary64 squares are exact and the test involves no rounding (a
// float x never equals such a square, which has more than 24 significant bits).
constexpr uint32_t kSweepPerThread = 64;
__global__ void sqrt_sweep_kernel(uint32_t first, uint32_t last, unsigned long long *fails, uint32_t *first_fail)
{
    const uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kSweepPerThread + first;
    unsigned long long bad = 0;
    uint32_t where = 0xffffffffu;
    for (uint32_t k = 0; k < kSweepPerThread; ++k) {
        const uint64_t u = i0 + k;
        if (u > last) break;
        const float x = __uint_as_float((uint32_t)u);
        const float s = sqrt_in_range(x);
        const uint32_t sb = __float_as_uint(s);
        const double sd = (double)s;
        const double lo = 0.5 * (sd + (double)__uint_as_float(sb - 1u));
        const double hi = 0.5 * (sd + (double)__uint_as_float(sb + 1u));
        const double xd = (double)x;
        const bool ok = sb > 0x00800000u && sb < 0x7f800000u && lo * lo < xd && xd < hi * hi;
        if (!ok) {
            ++bad;
            where = min(where, (uint32_t)u);
        }
    }
    if (bad) {
        atomicAdd(fails, bad);
        atomicMin(first_fail, where);
    }
}

__global__ void sweep_init_kernel(unsigned long long *fails, uint32_t *first_fail)
{
    if (threadIdx.x == 0) {
        *fails = 0ull;
        *first_fail = 0xffffffffu;
    }
}

uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_pack_met(const float *u0, const float *v0, const float *w0, const float *u1,
                           const float *v1, const float *w1, float4 *met, uint32_t npts,
                           hipStream_t st)
{
    pack_met_kernel<<<blocks_for(npts, 256), 256, 0, st>>>(u0, v0, w0, u1, v1, w1, met, npts);
    return hipGetLastError();
}

hipError_t launch_pack_pbl(const float *stage, float4 *pbl, uint32_t n, hipStream_t st)
{
    pack_pbl_kernel<<<blocks_for(n, 256), 256, 0, st>>>(stage, pbl, n);
    return hipGetLastError();
}

hipError_t launch_aos_to_soa(const fp_particle *aos, const uint32_t *slots, uint64_t n, Soa s,
                             hipStream_t st)
{
    if (n == 0) return hipSuccess;
    if (slots)
        aos_to_soa_kernel<<<blocks_for(n, 256), 256, 0, st>>>(
            reinterpret_cast<const uint32_t *>(aos), slots, n, s);
    else
        aos_to_soa_tiled<<<blocks_for(n, kT), kT, 0, st>>>(reinterpret_cast<const uint32_t *>(aos), n, s);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(Soa s, fp_particle *aos, uint64_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    soa_to_aos_tiled<<<blocks_for(n, kT), kT, 0, st>>>(s, reinterpret_cast<uint32_t *>(aos), n);
    return hipGetLastError();
}

uint64_t compact_blocks(uint32_t n) { return (n + kTile - 1) / kTile; }

hipError_t launch_compact(const Soa &in, const Soa &out, uint32_t n, uint32_t *tile_counts,
                          uint32_t *survivor_map, uint32_t *active_count, hipStream_t st)
{
    const uint32_t ntiles = (uint32_t)compact_blocks(n);
    if (ntiles == 0) return hipMemsetAsync(active_count, 0, sizeof(uint32_t), st);
    count_kernel<<<ntiles, kBlock, 0, st>>>(in, n, tile_counts);
    scan_kernel<<<1, 1024, 0, st>>>(tile_counts, ntiles, active_count);
    scatter_kernel<true><<<ntiles, kBlock, 0, st>>>(in, out, n, tile_counts, active_count, survivor_map);
    return hipGetLastError();
}

hipError_t launch_survivor_map(const Soa &in, uint32_t n, uint32_t *tile_counts, uint32_t *survivor_map,
                               uint32_t *active_count, hipStream_t st)
{
    const uint32_t ntiles = (uint32_t)compact_blocks(n);
    if (ntiles == 0) return hipMemsetAsync(active_count, 0, sizeof(uint32_t), st);
    count_kernel<<<ntiles, kBlock, 0, st>>>(in, n, tile_counts);
    scan_kernel<<<1, 1024, 0, st>>>(tile_counts, ntiles, active_count);
    scatter_kernel<false><<<ntiles, kBlock, 0, st>>>(in, in, n, tile_counts, active_count, survivor_map);
    return hipGetLastError();
}

namespace {
__global__ void store_zero_kernel(uint32_t *p) { if (threadIdx.x == 0) *p = 0u; }
}  // namespace

hipError_t launch_survivor_map_flags(const uint32_t *flags, uint32_t n, uint32_t *tile_counts,
                                     uint32_t *survivor_map, uint32_t *active_count, hipStream_t st)
{
    // a Soa whose every field is the flag array (capacity 0): active_at reads flags[i]
    Soa in;
    in.base = const_cast<uint32_t *>(flags);
    in.capacity = 0;
    const uint32_t ntiles = (uint32_t)compact_blocks(n);
    if (ntiles == 0) {
        store_zero_kernel<<<1, 64, 0, st>>>(active_count);
        return hipGetLastError();
    }
    count_kernel<<<ntiles, kBlock, 0, st>>>(in, n, tile_counts);
    scan_kernel<<<1, 1024, 0, st>>>(tile_counts, ntiles, active_count);
    scatter_kernel<false><<<ntiles, kBlock, 0, st>>>(in, in, n, tile_counts, active_count, survivor_map);
    return hipGetLastError();
}

hipError_t launch_math_probe(int fn, const float *x, const float *y, float *out, uint64_t n,
                             hipStream_t st)
{
    if (n == 0) return hipSuccess;
    math_probe_kernel<<<blocks_for(n, 256), 256, 0, st>>>(fn, x, y, out, n);
    return hipGetLastError();
}

hipError_t launch_level_probe(const float *table, uint32_t nz, uint32_t mode, const float *z, float *out, uint64_t n,
                              hipStream_t st)
{
    if (n == 0) return hipSuccess;
    level_probe_kernel<<<blocks_for(n, 256), 256, 0, st>>>(table, nz, mode, z, out, n);
    return hipGetLastError();
}

hipError_t launch_math_sweep(int fn, uint32_t first, uint32_t last, unsigned long long *fails,
                             uint32_t *first_fail, hipStream_t st)
{
    if (fn != FP_SWEEP_SQRT_STEP || first > last) return hipErrorInvalidValue;
    sweep_init_kernel<<<1, 64, 0, st>>>(fails, first_fail);
    const uint64_t n = (uint64_t)last - first + 1u, per_block = 256ull * kSweepPerThread;
    sqrt_sweep_kernel<<<(uint32_t)((n + per_block - 1) / per_block), 256, 0, st>>>(first, last, fails, first_fail);
    return hipGetLastError();
}

hipError_t launch_philox_blocks(const uint32_t counter[4], const uint32_t key[2], uint64_t n,
                                uint32_t *out, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    philox_kernel<<<blocks_for(n, 256), 256, 0, st>>>(counter[0], counter[1], counter[2],
                                                      counter[3], key[0], key[1], n, out);
    return hipGetLastError();
}

}  // namespace fp
