// fp_internal.hpp - device data layout, kernel argument blocks and the
// context object behind the C ABI (include/flexpart_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "flexpart_hip.h"
#include "fp_math.h"

namespace fp {

// SoA particle state: field f of slot i lives at base[f * capacity + i].  The
// field order is the reference's 96-byte Particle (particles/mod.rs:76-168),
// so an AoS <-> SoA transpose moves 32-bit word f of particle i to array f.
enum Field : uint32_t {
    F_POS_X = 0, F_POS_Y, F_POS_Z, F_CELL_X, F_CELL_Y, F_FLAGS,
    F_MASS0, F_MASS1, F_MASS2, F_MASS3,
    F_VEL_U, F_VEL_V, F_VEL_W, F_TURB_U, F_TURB_V, F_TURB_W,
    F_TIME, F_TIMESTEP, F_TIME_MEM, F_TIME_SPLIT, F_RELEASE_POINT, F_CLASS_ID, F_CBT, F_PAD0,
    F_COUNT
};
static_assert(F_COUNT * 4 == sizeof(fp_particle), "particle word count");

constexpr uint32_t kMaxLevels = 256;   // met levels held in LDS by the step kernel
// Particle slots per SoA set: the step kernels address a field word by a
// 32-bit byte offset from the field's base (slot << 2).
constexpr uint64_t kMaxSlots = 1ull << 30;
// Level table (device memory and LDS): the nz level heights, two +inf words
// (heights above the top, so that a lookup may read two levels past any
// level), then 1/w as a float, w = top height / kLevelBuckets, then for each
// height band [b w, (b+1) w) the first level k >= 1 whose height may be at or
// above a height in the band (uint16, two per word); level_of starts its
// search there.
constexpr uint32_t kLevelBuckets = 512;
constexpr uint32_t kLevelTableMax = kMaxLevels + 3 + kLevelBuckets / 2;
__host__ __device__ inline uint32_t level_table_words(uint32_t nz) { return nz + 3 + kLevelBuckets / 2; }
// StepArgs::has_heights: levels given as indices, as heights in metres, or
// as heights whose table admits the two-comparison lookup (a positive first
// height, at least two levels; level_table decides).
constexpr uint32_t kLevelIndex = 0, kLevelHeights = 1, kLevelHeightsBanded = 2;
constexpr uint32_t kWave = 64;         // CDNA wavefront width
constexpr uint32_t kXcds = 8;          // MI355X accelerator complex dies

// Pointers of one SoA set.
struct Soa {
    uint32_t *base = nullptr;  // F_COUNT * capacity words
    uint64_t capacity = 0;
    __host__ __device__ float *f(uint32_t field) const {
        return reinterpret_cast<float *>(base + (uint64_t)field * capacity);
    }
    __host__ __device__ int32_t *i(uint32_t field) const {
        return reinterpret_cast<int32_t *>(base + (uint64_t)field * capacity);
    }
    __host__ __device__ uint32_t *u(uint32_t field) const { return base + (uint64_t)field * capacity; }
    // The address of one field, recomputed from (base, capacity) at each use
    // with scalar instructions: the empty asm hides the capacity from loop-
    // invariant hoisting, so a kernel that touches many fields keeps two
    // scalar register pairs live instead of one pair per field (which spill
    // to vector-register lanes and cost a v_readlane per access).
    __host__ __device__ uint32_t *at(uint32_t field) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)capacity);
        uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(capacity >> 32));
        asm volatile("" : "+s"(lo), "+s"(hi));
        return base + (uint64_t)field * (((uint64_t)hi << 32) | lo);
#else
        return u(field);
#endif
    }
    __host__ __device__ float *atf(uint32_t field) const { return reinterpret_cast<float *>(at(field)); }
    __host__ __device__ int32_t *ati(uint32_t field) const { return reinterpret_cast<int32_t *>(at(field)); }
    // Word `slot` of a field addressed as the field's scalar base plus a
    // 32-bit byte offset (slot < 2^30), the form of a global access with a
    // scalar base register and a vector offset: no 64-bit vector address
    // arithmetic per access.
#if defined(__HIP_DEVICE_COMPILE__)
    template <class T> __device__ __attribute__((address_space(1))) T &w(uint32_t field, uint32_t slot) const
    {
        uint64_t p = reinterpret_cast<uint64_t>(at(field));
        asm volatile("" : "+s"(p));
        uint32_t off = slot << 2;
        asm("" : "+v"(off));
        using G = __attribute__((address_space(1))) char *;
        return *reinterpret_cast<__attribute__((address_space(1))) T *>(reinterpret_cast<G>(p) + off);
    }
    // The same word in a set of at most kSoa32Slots slots (at most 2^32
    // bytes): addressed from the set's base with the 32-bit byte offset
    // field * capacity * 4 + slot * 4, one scalar multiply and one vector add
    // per access instead of a 64-bit field address computed per access.
    template <class T> __device__ __attribute__((address_space(1))) T &w32(uint32_t field, uint32_t slot) const
    {
        const uint64_t p = reinterpret_cast<uint64_t>(base);
        const uint32_t stride = __builtin_amdgcn_readfirstlane((uint32_t)capacity << 2);
        // the slot's byte offset is opaque, so that the compiler adds the
        // field's offset to it (one add) rather than rescaling a slot sum
        uint32_t s4 = slot << 2;
        asm("" : "+v"(s4));
        uint32_t off = s4 + field * stride;
        asm("" : "+v"(off));
        using G = __attribute__((address_space(1))) char *;
        return *reinterpret_cast<__attribute__((address_space(1))) T *>(reinterpret_cast<G>(p) + off);
    }
#else
    template <class T> __host__ __device__ T &w(uint32_t field, uint32_t slot) const
    {
        return reinterpret_cast<T *>(u(field))[slot];
    }
#endif
};
// Capacity up to which a set's words are addressed with 32-bit byte offsets
// from its base (Soa::w32): F_COUNT * capacity * 4 <= 2^32.
constexpr uint64_t kSoa32Slots = (1ull << 32) / (F_COUNT * 4ull);

// Largest met grid (points) of a context, ensemble member or adjoint run:
// the packed two-time bracket (32 B per point) stays below 4 GiB, so a gather
// addresses it as its scalar base plus a 32-bit byte offset.
constexpr uint64_t kMaxMetPoints = 1ull << 27;

// Mass survival factors of one step, shared by the kernels (per-slot forcing)
// and the host (uniform forcing, evaluated once per step): the same float
// operations either way.
// dry_deposition.wgsl:70-106, for a particle below 2 href
__host__ __device__ inline float dry_survival(float vd, float dt, float href_in)
{
    const float href = fmaxf(href_in, 0.1f);
    const float ex = -fmaxf(vd, 0.0f) * fabsf(dt) / (2.0f * href);
    return fminf(fmaxf(fpm_expf(ex), 0.0f), 1.0f);
}
// wet_deposition.wgsl:73-111
__host__ __device__ inline float wet_survival(float lam, float frac, float dt_in)
{
    const float l = fmaxf(lam, 0.0f);
    const float dt = fabsf(dt_in);
    const float fr = fminf(fmaxf(frac, 0.0f), 1.0f);
    float prob = 0.0f;
    if (!(l <= 0.0f || dt <= 0.0f || fr <= 0.0f))
        prob = fminf(fmaxf((1.0f - fpm_expf(-l * dt)) * fr, 0.0f), 1.0f);
    return 1.0f - prob;
}

// The T at byte_off + imm bytes past a wavefront-uniform global base, loaded
// in the scalar-base form (base in scalar registers, 32-bit offset in a
// vector register, imm in the instruction: no 64-bit vector address
// arithmetic).
template <class T>
__host__ __device__ __forceinline__ T load_at(const T *base, uint32_t byte_off, uint32_t imm = 0)
{
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t p = reinterpret_cast<uint64_t>(base);
    asm volatile("" : "+s"(p));
    asm("" : "+v"(byte_off));
    using G = const __attribute__((address_space(1))) char *;
    return *reinterpret_cast<const __attribute__((address_space(1))) T *>(reinterpret_cast<G>(p) + byte_off + imm);
#else
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off + imm);
#endif
}

// Concentration-grid configuration shared by the fused and standalone paths.
struct GridArgs {
    uint32_t *counts;
    uint32_t *mass_fp;
    uint32_t nx, ny, nz;
    uint32_t refine_x, refine_y;
    uint32_t species;
    float mass_scale;
    uint32_t metres;  // outheights[nz-1] > 0
    // the fused step's cells (met cells cx < met nx, cy < met ny) refine into
    // int32 indices (met nx * refine_x and met ny * refine_y below 2^31) and
    // the output heights, if used, are non-decreasing: grid_cell<true>
    uint32_t fast;
    float outheights[FP_MAX_OUTPUT_LEVELS];
};

// Where a gathered step (launch_step_gathered) takes each particle's step.
constexpr int kGatherNone = 0, kGatherToSlot = 1, kGatherFromSlot = 2;

// Everything the fused step kernel reads, passed by value.
struct StepArgs {
    Soa s;
    uint32_t n;             // dispatch count
    uint32_t n_species;
    // met: packed 32-byte records {u0,v0,w0,u1,v1,w1,0,0} per grid point
    const float4 *met;
    uint32_t nx, ny, nz;
    const float *level_heights;  // nz floats (device)
    uint32_t has_heights;
    float zmax;
    // PBL: two 16-byte records per cell (i, j): {ustar, wstar, hmix, oli} and
    // {L, unstable sigma_u, regime, 0} (pbl_cell_terms_dev); precip optional
    const float4 *pbl;
    const float *precip;
    // forcing
    const float *dry_vd_slots, *wet_lambda_slots, *wet_fraction_slots;
    // scalars
    float dt_adv, dt, alpha, x_scale, y_scale, z_scale;
    uint32_t key0, key1, c0, c1, c2, c3;
    uint32_t n_substeps, blocks_per_particle;
    float min_height, rho_grad;
    uint32_t dry_mode;
    float dry_vd, dry_href;
    uint32_t wet_mode;
    float wet_lambda, wet_fraction, wet_a, wet_b;
    // survival factors of the uniform dry and wet modes (dry_survival and
    // wet_survival of the scalars above, evaluated on the host)
    float dry_sv, wet_sv;
    uint32_t loss_mode;
    float loss_min_mass;
    uint32_t grid_fused;
    GridArgs grid;
    uint32_t ppt;           // particles per thread
    double log_1em3;        // fpm_log_d(1e-3f): the log of the unstable profile's zeta floor
    // launch-uniform values the step would otherwise recompute in every lane
    // (fill_step_scalars, with the same f32 operations): (float)(n - 1) per
    // grid axis, the sub-step length dt / n_substeps and the dry-deposition
    // height cut 2 max(href, 0.1)
    float nxm, nym, nzm, dt_sub, dry_zcut;
    // the launch's part of the range check for unscaled quotients in the
    // turbulence (lean_lane): dt and dt_sub in [2^-20, 2^20], z_min in [2^-40, 1)
    uint32_t lean_div;
    // gathered steps (launch_step_gathered): slot i of s receives the words
    // the step writes of src slot gather_map[i] advanced by one step, i.e. the
    // step of the state that a compaction with survivor map gather_map would
    // produce; the other words of s are left as they are
    const uint32_t *gather_map;
    Soa src;
};

// Ensemble launch tables (device memory).
struct EnsMember {
    StepArgs a;              // member's particle range, met, grid, key, scalars
    uint32_t base_ctr[4];    // Philox counter at clock step 0
    uint64_t ctr_per_step;   // dispatch * blocks per particle
};
struct EnsBlock { uint32_t member; uint32_t first; };  // first member-local slot
struct EnsClock {
    uint32_t next_step, step;
    float alpha;
    uint32_t pad;
    double t0_in_bracket, dt_seconds, bracket_seconds;
};

// Which of the two SoA sets holds the current data of each slot range.
// Compaction and sort write the dispatch range into the other set and swap;
// slots beyond the dispatch range keep their data in the set that last wrote
// them and are copied over lazily when a download, an upload or a wider
// dispatch range needs them.
struct Segments {
    struct Seg { uint64_t a, b; int owner; };
    std::vector<Seg> v;
    void reset(uint64_t cap, int owner) { v.assign(1, Seg{0, cap, owner}); }
    void assign(uint64_t a, uint64_t b, int owner)
    {
        if (a >= b) return;
        std::vector<Seg> out;
        for (const Seg &s : v) {
            if (s.b <= a || s.a >= b) { out.push_back(s); continue; }
            if (s.a < a) out.push_back(Seg{s.a, a, s.owner});
            if (s.b > b) out.push_back(Seg{b, s.b, s.owner});
        }
        out.push_back(Seg{a, b, owner});
        std::sort(out.begin(), out.end(), [](const Seg &x, const Seg &y) { return x.a < y.a; });
        v.clear();
        for (const Seg &s : out) {
            if (!v.empty() && v.back().owner == s.owner && v.back().b == s.a) v.back().b = s.b;
            else v.push_back(s);
        }
    }
};

}  // namespace fp

namespace fp {
// Device work buffers of a Morton sort of up to `cap` slots (sort_morton_range).
struct SortWork {
    double *bounds_partial = nullptr;  // kSortPartials * 7
    double *bounds_dev = nullptr;      // 8
    void *keys = nullptr, *keys_alt = nullptr;        // cap u64 keys each
    uint32_t *vals = nullptr, *vals_alt = nullptr;    // cap slots each; vals_alt = new_to_old after a sort
    void *temp = nullptr;                             // rocPRIM temporary storage, grown on demand
    size_t temp_bytes = 0;
    uint64_t cap = 0;
};
constexpr uint32_t kSortPartials = 2048;
hipError_t sort_work_alloc(SortWork &w, uint64_t cap);
void sort_work_free(SortWork &w);
}  // namespace fp

struct fp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    fp_ctx_desc desc{};
    uint64_t dispatch = 0;

    fp::Soa soa[2];
    int cur = 0;
    fp::Segments segs;

    // met
    float4 *met = nullptr;         // 2 float4 per point
    float *level_heights = nullptr;
    std::vector<float> level_heights_host;
    bool met_ready = false;
    uint32_t has_heights = fp::kLevelIndex;  // StepArgs::has_heights
    float zmax = 0.0f;
    float *met_stage = nullptr;    // 6 * npoints floats for host uploads

    // PBL (2 records per cell)
    float4 *pbl = nullptr;
    float *precip = nullptr;
    bool pbl_ready = false;
    bool precip_ready = false;
    float *pbl_stage = nullptr;    // 4 * nx*ny

    // surface fields of the two bracket ends (FP_SURF_COUNT * nx*ny floats each)
    float *surf[2] = {nullptr, nullptr};
    int64_t surf_time[2] = {0, 0};
    bool surf_ready[2] = {false, false};

    // forcing
    float *dry_vd = nullptr, *wet_lambda = nullptr, *wet_fraction = nullptr;

    // grid
    uint32_t *grid_counts = nullptr, *grid_mass = nullptr;
    fp_grid_params grid_params{};
    bool grid_configured = false;

    // staging for AoS exchange
    fp_particle *aos_stage = nullptr;
    uint64_t aos_stage_cap = 0;
    uint32_t *slot_stage = nullptr;
    uint64_t slot_stage_cap = 0;

    // compaction
    uint32_t *block_counts = nullptr;
    uint64_t block_counts_cap = 0;
    uint32_t *survivor_map = nullptr;
    uint32_t *active_count_dev = nullptr;
    uint32_t *active_count_host = nullptr;  // pinned

    // sort
    fp::SortWork sortw;

    // asynchronous bracket prefetch
    hipStream_t copy_stream = nullptr;
    hipEvent_t prefetch_done = nullptr, stage_free = nullptr;
    float *prefetch_stage = nullptr;  // 6 * npts
    bool prefetch_pending = false, stage_used = false;
    std::vector<float> prefetch_lh;

    // hybrid-level to height transform buffers
    float *vt_buf = nullptr;  // ps (ncol) | T | q | p | z (ncell each) | ak | bk | mean (nz each)
    unsigned long long *vt_err = nullptr;

    // concentration output: interval sums and finalize buffers
    unsigned long long *out_acc_count = nullptr, *out_acc_mass = nullptr;
    uint32_t out_samples = 0;
    uint32_t *out_count = nullptr;
    float *out_mass = nullptr, *out_conc = nullptr, *out_vol = nullptr;

    // source-term injection plans
    void *release_jobs = nullptr;
    uint64_t release_jobs_bytes = 0;

    // launch shape
    uint32_t block_size = 256;
    uint32_t ppt = 1;

    // profiling
    bool profiling = false;
    struct Ev { hipEvent_t a, b; int kernel; };
    std::vector<Ev> events;
    std::vector<Ev> free_events;
    double prof_ms[FP_KERNEL_COUNT] = {};
    uint64_t prof_n[FP_KERNEL_COUNT] = {};

    const fp::Soa &s() const { return soa[cur]; }
};

namespace fp {

// Kernel launchers (fp_step.hip, fp_aux.hip).
hipError_t launch_step(const StepArgs &a, uint32_t block, hipStream_t st);
// one step of a.n slots through a.gather_map (block 256, a.ppt rows); the
// particle steps at its destination slot, or at its source slot
hipError_t launch_step_gathered(const StepArgs &a, bool at_source_slot, hipStream_t st);
// one step of slots [0, a.n) of a.src into the same slots of a.s (stepped words)
hipError_t launch_step_out_of_place(const StepArgs &a, hipStream_t st);
// ppt: particles per thread of every member (a.ppt), selecting the kernel
// variant; soa32: the ensemble's particle set holds at most kSoa32Slots slots
hipError_t launch_ens_step(const EnsMember *members, const EnsBlock *blocks, EnsClock *clock,
                           uint32_t nblocks, uint32_t block, uint32_t ppt, bool sub4, bool soa32, hipStream_t st);
hipError_t launch_pack_met(const float *u0, const float *v0, const float *w0, const float *u1,
                           const float *v1, const float *w1, float4 *met, uint32_t npts,
                           hipStream_t st);
hipError_t launch_pack_pbl(const float *stage, float4 *pbl, uint32_t n, hipStream_t st);
hipError_t launch_aos_to_soa(const fp_particle *aos, const uint32_t *slots, uint64_t n, Soa s,
                             hipStream_t st);
hipError_t launch_soa_to_aos(Soa s, fp_particle *aos, uint64_t n, hipStream_t st);
hipError_t launch_grid(const Soa &s, uint32_t n, const GridArgs &g, hipStream_t st);
hipError_t launch_compact(const Soa &in, const Soa &out, uint32_t n, uint32_t *block_counts,
                          uint32_t *survivor_map, uint32_t *active_count, hipStream_t st);
// the survivor map and active count of launch_compact without moving the records
hipError_t launch_survivor_map(const Soa &in, uint32_t n, uint32_t *block_counts, uint32_t *survivor_map,
                               uint32_t *active_count, hipStream_t st);
uint64_t compact_blocks(uint32_t n);
// launch_survivor_map over a plain flag array (bit 0 = active) with the
// caller's tile-count buffer (compact_blocks(n) words); active_count is
// written on the device (zero for n = 0, by a kernel: no memset node)
hipError_t launch_survivor_map_flags(const uint32_t *flags, uint32_t n, uint32_t *tile_counts,
                                     uint32_t *survivor_map, uint32_t *active_count, hipStream_t st);
hipError_t launch_philox_blocks(const uint32_t counter[4], const uint32_t key[2], uint64_t n,
                                uint32_t *out, hipStream_t st);
hipError_t launch_math_probe(int fn, const float *x, const float *y, float *out, uint64_t n,
                             hipStream_t st);
// level_of of n heights z over a level table (level_table_words(nz) words).
hipError_t launch_level_probe(const float *table, uint32_t nz, uint32_t mode, const float *z, float *out, uint64_t n,
                              hipStream_t st);
hipError_t launch_math_sweep(int fn, uint32_t first, uint32_t last, unsigned long long *fails,
                             uint32_t *first_fail, hipStream_t st);
// Source-term injection (fp_release.hip).
hipError_t release_plan_and_launch(const fp_grid_domain &g, const fp_release *rel, uint32_t n_rel,
                                   uint64_t *released, int64_t now, int32_t time, const uint32_t *slots_dev,
                                   uint64_t first_slot, uint64_t max_particles, const Soa &s, void **jobs_dev,
                                   uint64_t *jobs_cap_bytes, hipStream_t st, uint64_t *injected, int *status);
// PBL records from the two surface bracket ends (fp_surface.hip).
hipError_t launch_pbl_surface(const float *const s0[FP_SURF_COUNT], const float *const s1[FP_SURF_COUNT], float alpha,
                              const fp_pbl_options &o, float4 *pbl, float *pass, uint32_t n, hipStream_t st);
// Hybrid-level to height transform (fp_met.hip); err: two u64 atomic-min keys.
hipError_t launch_vertical_transform(const float *ps, const float *t, const float *q, const float *ak,
                                     const float *bk, float *pressure, float *height, float *mean,
                                     unsigned long long *err, uint32_t ncol, uint32_t nz, hipStream_t st);
// Concentration output (fp_output.hip).
hipError_t launch_output_sample(const uint32_t *counts, const uint32_t *mass, unsigned long long *acc_count,
                                unsigned long long *acc_mass, uint64_t n, hipStream_t st);
hipError_t launch_output_finalize(const unsigned long long *acc_count, const unsigned long long *acc_mass,
                                  uint32_t n_samples, float mass_scale, const float *vol, uint32_t ny, uint32_t nz,
                                  uint32_t *count_out, float *mass_out, float *conc_out, uint64_t n, hipStream_t st);
void output_volumes(const fp_output_geometry &g, const float *outheights, uint32_t ny, uint32_t nz, float *vol);
int write_netcdf(const char *path, const fp_netcdf_snapshot &s, int version, std::string &err);
// Adjoint of one step (fp_adj.hip): lam (7 x stride floats: x, y, z, u', v',
// w', m0 of the step's output) is replaced by the adjoint of its input; au/av
// (nx*ny doubles, may be null) accumulate the adjoint of the horizontal wind.
hipError_t launch_adj_step(const StepArgs &a, float *lam, uint64_t stride, double *au, double *av, hipStream_t st);
// Morton sort: keys + stable radix sort + gather (fp_sort.hip).
// Morton sort of slots [0, n) of `in` into `out` (which may be `in`: the
// records are gathered from an AoS copy in aos_stage, >= n records).
hipError_t sort_morton_range(SortWork &w, const fp_sort_params &p, const Soa &in, const Soa &out,
                             fp_particle *aos_stage, uint32_t n, hipStream_t st);

// Sets the message fp_last_error(NULL) returns (no allocation; fp_api.hip).
void set_context_free_error(const char *msg);

// The level table of nz non-decreasing heights (level-index mode when the top
// height is not positive), level_table_words(nz) words; returns the
// StepArgs::has_heights mode the table is read in (fp_api.hip).
uint32_t level_table(const float *lh, uint32_t nz, std::vector<float> &out);

// Shared by the single-run and ensemble entry points (fp_api.hip).
// Scalar step arguments (timestep, scales, Philox key and counter, sub-steps,
// deposition and loss modes) from the public parameter block.
void fill_step_scalars(StepArgs &a, const fp_step_params &p);
GridArgs make_grid_args(uint32_t *counts, uint32_t *mass_fp, uint32_t nx, uint32_t ny, uint32_t nz,
                        const fp_grid_params &gp,
                        uint32_t met_nx = 0, uint32_t met_ny = 0);
// 128-bit little-endian counter + off (gpu/langevin_fused.rs:229-233).
void counter_advance(const uint32_t c[4], uint64_t off, uint32_t out[4]);
// Philox counter blocks one step consumes: dispatch * blocks per particle.
inline uint64_t counter_per_step(uint64_t dispatch, uint32_t n_substeps)
{
    return dispatch * (n_substeps <= 2 ? 1ull : 2ull);
}

}  // namespace fp
