// fp_api.hip - the C ABI (include/flexpart_hip.h) over the gfx950 kernels.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "fp_internal.hpp"
#include "fp_math.h"

static_assert(sizeof(fp_particle) == 96, "fp_particle must match the reference Particle");
static_assert(sizeof(fp_step_params) == 132, "fp_step_params layout is part of the ABI");
static_assert(sizeof(fp_ctx_desc) == 64, "fp_ctx_desc layout is part of the ABI");
static_assert(sizeof(fp_grid_params) == 80, "fp_grid_params layout is part of the ABI");
static_assert(sizeof(fp_sort_params) == 56, "fp_sort_params layout is part of the ABI");
static_assert(sizeof(fp_autotune_params) == 64, "fp_autotune_params layout is part of the ABI");
static_assert(sizeof(fp_autotune_result) == 32, "fp_autotune_result layout is part of the ABI");
static_assert(sizeof(fp_grid_domain) == 40, "fp_grid_domain layout is part of the ABI");
static_assert(sizeof(fp_release) == 88, "fp_release layout is part of the ABI");

namespace {

// Message of the last failed call that has no context (fp_ctx_create and the
// context-free entry points), per thread; a fixed buffer, so that setting it
// allocates nothing.
thread_local char g_create_error[512];

void set_create_error(const char *msg)
{
    std::snprintf(g_create_error, sizeof g_create_error, "%s", msg);
}

int fail(fp_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg; else set_create_error(msg.c_str());
    return code;
}

#define FP_HIP(ctx, expr)                                                                     \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((ctx), e_ == hipErrorOutOfMemory ? FP_E_NOMEM : FP_E_DEVICE,          \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                   \
    } while (0)

#define FP_CHECK(ctx, cond, msg)                                                              \
    do {                                                                                      \
        if (!(cond)) return fail((ctx), FP_E_INVALID, (msg));                                 \
    } while (0)

template <typename T>
hipError_t dalloc(T **p, uint64_t count)
{
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

int ensure_current(fp_ctx *ctx, uint64_t a, uint64_t b)
{
    b = std::min<uint64_t>(b, ctx->desc.capacity);
    if (a >= b) return FP_OK;
    std::vector<fp::Segments::Seg> todo;
    for (const auto &s : ctx->segs.v)
        if (s.owner != ctx->cur && s.b > a && s.a < b)
            todo.push_back(fp::Segments::Seg{std::max(s.a, a), std::min(s.b, b), s.owner});
    for (const auto &s : todo) {
        const fp::Soa &src = ctx->soa[s.owner];
        const fp::Soa &dst = ctx->soa[ctx->cur];
        for (uint32_t f = 0; f < fp::F_COUNT; ++f)
            FP_HIP(ctx, hipMemcpyAsync(dst.u(f) + s.a, src.u(f) + s.a, (s.b - s.a) * 4,
                                       hipMemcpyDeviceToDevice, ctx->stream));
        ctx->segs.assign(s.a, s.b, ctx->cur);
    }
    return FP_OK;
}

int ensure_stage(fp_ctx *ctx, uint64_t n)
{
    if (n <= ctx->aos_stage_cap) return FP_OK;
    if (ctx->aos_stage) (void)hipFree(ctx->aos_stage);
    ctx->aos_stage = nullptr;
    ctx->aos_stage_cap = 0;
    FP_HIP(ctx, dalloc(&ctx->aos_stage, n));
    ctx->aos_stage_cap = n;
    return FP_OK;
}

int ensure_slot_stage(fp_ctx *ctx, uint64_t n)
{
    if (n <= ctx->slot_stage_cap) return FP_OK;
    if (ctx->slot_stage) (void)hipFree(ctx->slot_stage);
    ctx->slot_stage = nullptr;
    ctx->slot_stage_cap = 0;
    FP_HIP(ctx, dalloc(&ctx->slot_stage, n));
    ctx->slot_stage_cap = n;
    return FP_OK;
}

void prof_begin(fp_ctx *ctx, fp_ctx::Ev &ev, int kernel)
{
    ev.kernel = kernel;
    if (!ctx->profiling) return;
    if (!ctx->free_events.empty()) {
        ev = ctx->free_events.back();
        ctx->free_events.pop_back();
        ev.kernel = kernel;
    } else {
        (void)hipEventCreate(&ev.a);
        (void)hipEventCreate(&ev.b);
    }
    (void)hipEventRecord(ev.a, ctx->stream);
}

void prof_end(fp_ctx *ctx, fp_ctx::Ev &ev)
{
    if (!ctx->profiling) return;
    (void)hipEventRecord(ev.b, ctx->stream);
    ctx->events.push_back(ev);
}

int check_levels(fp_ctx *ctx, const float *lh)
{
    const uint32_t nz = ctx->desc.met_nz;
    for (uint32_t k = 0; k < nz; ++k)
        FP_CHECK(ctx, std::isfinite(lh[k]), "level heights must be finite");
    if (lh[nz - 1] > 0.0f)
        for (uint32_t k = 1; k < nz; ++k)
            FP_CHECK(ctx, lh[k] >= lh[k - 1], "level heights must be non-decreasing");
    return FP_OK;
}

int validate_levels(fp_ctx *ctx, const float *lh)
{
    if (int rc = check_levels(ctx, lh)) return rc;
    const uint32_t nz = ctx->desc.met_nz;
    const bool metres = lh[nz - 1] > 0.0f;
    ctx->has_heights = fp::level_table(lh, nz, ctx->level_heights_host);
    ctx->zmax = metres ? lh[nz - 1] : (float)(nz - 1);
    return FP_OK;
}

void destroy(fp_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    auto fr = [](void *p) { if (p) (void)hipFree(p); };
    fr(ctx->soa[0].base); fr(ctx->soa[1].base);
    fr(ctx->met); fr(ctx->level_heights); fr(ctx->met_stage);
    fr(ctx->pbl); fr(ctx->precip); fr(ctx->pbl_stage); fr(ctx->surf[0]); fr(ctx->surf[1]);
    fr(ctx->out_acc_count); fr(ctx->out_acc_mass); fr(ctx->out_count); fr(ctx->out_mass); fr(ctx->out_conc);
    fr(ctx->out_vol); fr(ctx->prefetch_stage); fr(ctx->vt_buf); fr(ctx->vt_err);
    if (ctx->copy_stream) { (void)hipStreamSynchronize(ctx->copy_stream); (void)hipStreamDestroy(ctx->copy_stream); }
    if (ctx->prefetch_done) (void)hipEventDestroy(ctx->prefetch_done);
    if (ctx->stage_free) (void)hipEventDestroy(ctx->stage_free);
    fr(ctx->dry_vd); fr(ctx->wet_lambda); fr(ctx->wet_fraction);
    fr(ctx->grid_counts); fr(ctx->grid_mass);
    fr(ctx->aos_stage); fr(ctx->slot_stage);
    fr(ctx->block_counts); fr(ctx->survivor_map); fr(ctx->active_count_dev);
    fp::sort_work_free(ctx->sortw);
    fr(ctx->release_jobs);
    if (ctx->active_count_host) (void)hipHostFree(ctx->active_count_host);
    for (auto &ev : ctx->events) { (void)hipEventDestroy(ev.a); (void)hipEventDestroy(ev.b); }
    for (auto &ev : ctx->free_events) { (void)hipEventDestroy(ev.a); (void)hipEventDestroy(ev.b); }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // namespace

// ---------------------------------------------------------------- shared helpers

fp::GridArgs fp::make_grid_args(uint32_t *counts, uint32_t *mass_fp, uint32_t nx, uint32_t ny, uint32_t nz,
                                const fp_grid_params &gp, uint32_t met_nx, uint32_t met_ny)
{
    fp::GridArgs g{};
    g.counts = counts;
    g.mass_fp = mass_fp;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.refine_x = gp.refine_x;
    g.refine_y = gp.refine_y;
    g.species = gp.species;
    g.mass_scale = gp.mass_scale;
    g.metres = nz ? (gp.outheights[nz - 1] > 0.0f) : 0u;
    for (int k = 0; k < FP_MAX_OUTPUT_LEVELS; ++k) g.outheights[k] = gp.outheights[k];
    bool sorted = true;
    for (uint32_t k = 1; k < nz; ++k) sorted = sorted && gp.outheights[k] >= gp.outheights[k - 1];
    const uint64_t lim = 0x7fffffffull;
    g.fast = met_nx && met_ny && (uint64_t)met_nx * gp.refine_x <= lim && (uint64_t)met_ny * gp.refine_y <= lim &&
             nz >= 1 && (!g.metres || sorted);
    return g;
}

void fp::fill_step_scalars(fp::StepArgs &a, const fp_step_params &p)
{
    a.dt = p.dt_seconds;
    a.dt_adv = p.direction < 0 ? -p.dt_seconds : p.dt_seconds;
    a.alpha = std::min(std::max(p.alpha, 0.0f), 1.0f);  // gpu/advection.rs:639
    a.x_scale = p.x_scale;
    a.y_scale = p.y_scale;
    a.z_scale = p.z_scale;
    a.key0 = p.key[0];
    a.key1 = p.key[1];
    a.c0 = p.counter[0];
    a.c1 = p.counter[1];
    a.c2 = p.counter[2];
    a.c3 = p.counter[3];
    a.n_substeps = p.n_substeps;
    const uint32_t nv = std::max(p.n_substeps, 1u);
    a.blocks_per_particle = (2u + nv + 3u) / 4u;  // langevin_fused.wgsl:435-436
    a.min_height = p.min_height_m;
    a.rho_grad = p.rho_grad_over_rho;
    a.dry_mode = p.dry_mode;
    a.dry_vd = p.dry_vd;
    a.dry_href = p.dry_href;
    a.wet_mode = p.wet_mode;
    a.wet_lambda = p.wet_lambda;
    a.wet_fraction = p.wet_fraction;
    a.wet_a = p.wet_a;
    a.wet_b = p.wet_b;
    a.dry_sv = fp::dry_survival(p.dry_vd, p.dt_seconds, p.dry_href);
    a.wet_sv = fp::wet_survival(p.wet_lambda, p.wet_fraction, p.dt_seconds);
    a.loss_mode = p.loss_mode;
    a.loss_min_mass = p.loss_min_mass;
    a.grid_fused = p.grid_fused ? 1u : 0u;
    a.log_1em3 = fpm_log_d((double)1.0e-3f);
    // derived from the grid dimensions, which every caller sets first
    a.nxm = (float)(a.nx - 1u);
    a.nym = (float)(a.ny - 1u);
    a.nzm = (float)(a.nz - 1u);
    a.dt_sub = a.dt / (float)nv;
    a.dry_zcut = 2.0f * std::fmax(a.dry_href, 0.1f);  // fmaxf semantics: 0.1 for a NaN href
    auto in = [](float v, float lo, float hi) { return v >= lo && v <= hi; };
    a.lean_div = in(a.dt, 0x1p-20f, 0x1p20f) && in(a.dt_sub, 0x1p-20f, 0x1p20f) && a.min_height >= 0x1p-40f &&
                 a.min_height < 1.0f;
}

void fp::counter_advance(const uint32_t c[4], uint64_t off, uint32_t out[4])
{
    const uint64_t lo = (uint64_t)c[0] + (uint32_t)off;
    const uint64_t mid = (uint64_t)c[1] + (uint32_t)(off >> 32) + (lo >> 32);
    const uint64_t hi = (uint64_t)c[2] + (mid >> 32);
    out[0] = (uint32_t)lo;
    out[1] = (uint32_t)mid;
    out[2] = (uint32_t)hi;
    out[3] = c[3] + (uint32_t)(hi >> 32);
}

extern "C" {

const char *fp_version(void) { return "flexpart-hip 0.1 gfx950"; }

const char *fp_last_error(const fp_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error;
}

int fp_ctx_create(int device_ordinal, const fp_ctx_desc *desc, fp_ctx **out)
{
    if (!out || !desc) return fail(nullptr, FP_E_INVALID, "null argument");
    *out = nullptr;
    const fp_ctx_desc &d = *desc;
    if (d.capacity == 0 || d.capacity > fp::kMaxSlots)
        return fail(nullptr, FP_E_INVALID, "capacity must be in [1, 2^30]");
    if (d.met_nx == 0 || d.met_ny == 0 || d.met_nz == 0)
        return fail(nullptr, FP_E_INVALID, "met grid dimensions must be non-zero");
    if (d.met_nz > fp::kMaxLevels) return fail(nullptr, FP_E_INVALID, "met_nz exceeds 256");
    const uint64_t npts = (uint64_t)d.met_nx * d.met_ny * d.met_nz;
    if (npts > fp::kMaxMetPoints) return fail(nullptr, FP_E_INVALID, "met grid too large (at most 2^27 points)");
    if (d.n_species < 1 || d.n_species > FP_MAX_SPECIES)
        return fail(nullptr, FP_E_INVALID, "n_species must be 1..4");
    const uint64_t ncell = (uint64_t)d.out_nx * d.out_ny * d.out_nz;
    if (ncell >= 0x80000000ull) return fail(nullptr, FP_E_INVALID, "output grid too large");
    if (ncell && d.out_nz > FP_MAX_OUTPUT_LEVELS)
        return fail(nullptr, FP_E_INVALID, "out_nz exceeds 16");

    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(nullptr, FP_E_DEVICE, "no HIP device");
    if (device_ordinal < 0 || device_ordinal >= ndev)
        return fail(nullptr, FP_E_INVALID, "device ordinal out of range");
    e = hipSetDevice(device_ordinal);
    if (e != hipSuccess) return fail(nullptr, FP_E_DEVICE, hipGetErrorString(e));

    fp_ctx *ctx = new (std::nothrow) fp_ctx();
    if (!ctx) return fail(nullptr, FP_E_NOMEM, "host allocation failed");
    ctx->device = device_ordinal;
    ctx->desc = d;
    ctx->dispatch = d.capacity;
    auto bail = [&](int code) {
        set_create_error(ctx->err.c_str());
        destroy(ctx);
        return code;
    };
#define CREATE_HIP(expr)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                     \
            return bail(e_ == hipErrorOutOfMemory ? FP_E_NOMEM : FP_E_DEVICE);                \
        }                                                                                     \
    } while (0)
    CREATE_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
        ctx->soa[k].capacity = d.capacity;
        CREATE_HIP(dalloc(&ctx->soa[k].base, d.capacity * fp::F_COUNT));
    }
    CREATE_HIP(hipMemsetAsync(ctx->soa[0].base, 0, d.capacity * fp::F_COUNT * 4, ctx->stream));
    CREATE_HIP(dalloc(&ctx->met, npts * 2));
    CREATE_HIP(dalloc(&ctx->level_heights, (uint64_t)fp::level_table_words(d.met_nz)));
    const uint64_t np = (uint64_t)d.met_nx * d.met_ny;
    CREATE_HIP(dalloc(&ctx->pbl, np * 2));
    CREATE_HIP(dalloc(&ctx->precip, np));
    if (ncell) {
        CREATE_HIP(dalloc(&ctx->grid_counts, ncell));
        CREATE_HIP(dalloc(&ctx->grid_mass, ncell));
        CREATE_HIP(hipMemsetAsync(ctx->grid_counts, 0, ncell * 4, ctx->stream));
        CREATE_HIP(hipMemsetAsync(ctx->grid_mass, 0, ncell * 4, ctx->stream));
    }
    CREATE_HIP(dalloc(&ctx->block_counts, fp::compact_blocks((uint32_t)d.capacity) + 1));
    CREATE_HIP(dalloc(&ctx->survivor_map, d.capacity));
    CREATE_HIP(dalloc(&ctx->active_count_dev, 1));
    CREATE_HIP(hipHostMalloc(reinterpret_cast<void **>(&ctx->active_count_host), sizeof(uint32_t)));
    CREATE_HIP(fp::sort_work_alloc(ctx->sortw, d.capacity));
    CREATE_HIP(hipStreamSynchronize(ctx->stream));
#undef CREATE_HIP
    ctx->segs.reset(d.capacity, 0);
    ctx->grid_params.refine_x = 1;
    ctx->grid_params.refine_y = 1;
    ctx->grid_params.mass_scale = 1.0e6f;
    *out = ctx;
    return FP_OK;
}

int fp_ctx_destroy(fp_ctx *ctx)
{
    destroy(ctx);
    return FP_OK;
}

int fp_sync(fp_ctx *ctx)
{
    FP_CHECK(ctx, ctx, "null context");
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

// ---------------------------------------------------------------- particles

static int check_species(fp_ctx *ctx, const fp_particle *p, uint64_t n)
{
    for (uint64_t r = 0; r < n; ++r)
        for (uint32_t s = ctx->desc.n_species; s < FP_MAX_SPECIES; ++s)
            if (p[r].mass[s] != 0.0f)
                return fail(ctx, FP_E_INVALID,
                            "mass slot beyond n_species is non-zero (slot record " +
                                std::to_string(r) + ")");
    return FP_OK;
}

int fp_particles_upload(fp_ctx *ctx, const fp_particle *p, uint64_t n)
{
    FP_CHECK(ctx, ctx && (p || n == 0), "null argument");
    FP_CHECK(ctx, n <= ctx->desc.capacity, "upload exceeds capacity");
    if (int rc = check_species(ctx, p, n)) return rc;
    if (int rc = ensure_stage(ctx, n)) return rc;
    FP_HIP(ctx, hipMemcpyAsync(ctx->aos_stage, p, n * sizeof(fp_particle), hipMemcpyHostToDevice,
                               ctx->stream));
    FP_HIP(ctx, fp::launch_aos_to_soa(ctx->aos_stage, nullptr, n, ctx->s(), ctx->stream));
    ctx->segs.assign(0, n, ctx->cur);
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));  // p may be released by the caller
    return FP_OK;
}

int fp_particles_upload_slots(fp_ctx *ctx, const uint32_t *slots, const fp_particle *p, uint64_t n)
{
    FP_CHECK(ctx, ctx && ((slots && p) || n == 0), "null argument");
    if (n == 0) return FP_OK;
    uint64_t lo = UINT64_MAX, hi = 0;
    for (uint64_t r = 0; r < n; ++r) {
        FP_CHECK(ctx, slots[r] < ctx->desc.capacity, "slot index out of range");
        lo = std::min<uint64_t>(lo, slots[r]);
        hi = std::max<uint64_t>(hi, slots[r]);
    }
    if (int rc = check_species(ctx, p, n)) return rc;
    if (int rc = ensure_current(ctx, lo, hi + 1)) return rc;
    if (int rc = ensure_stage(ctx, n)) return rc;
    if (int rc = ensure_slot_stage(ctx, n)) return rc;
    FP_HIP(ctx, hipMemcpyAsync(ctx->aos_stage, p, n * sizeof(fp_particle), hipMemcpyHostToDevice,
                               ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(ctx->slot_stage, slots, n * 4, hipMemcpyHostToDevice, ctx->stream));
    FP_HIP(ctx, fp::launch_aos_to_soa(ctx->aos_stage, ctx->slot_stage, n, ctx->s(), ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

int fp_particles_download(fp_ctx *ctx, fp_particle *out, uint64_t n)
{
    FP_CHECK(ctx, ctx && (out || n == 0), "null argument");
    FP_CHECK(ctx, n <= ctx->desc.capacity, "download exceeds capacity");
    if (n == 0) return FP_OK;
    if (int rc = ensure_current(ctx, 0, n)) return rc;
    if (int rc = ensure_stage(ctx, n)) return rc;
    FP_HIP(ctx, fp::launch_soa_to_aos(ctx->s(), ctx->aos_stage, n, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(out, ctx->aos_stage, n * sizeof(fp_particle), hipMemcpyDeviceToHost,
                               ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

int fp_particles_set_dispatch_count(fp_ctx *ctx, uint64_t n)
{
    FP_CHECK(ctx, ctx, "null context");
    FP_CHECK(ctx, n <= ctx->desc.capacity, "dispatch count exceeds capacity");
    if (int rc = ensure_current(ctx, 0, n)) return rc;
    ctx->dispatch = n;
    return FP_OK;
}

uint64_t fp_particles_dispatch_count(const fp_ctx *ctx) { return ctx ? ctx->dispatch : 0; }

int fp_particles_soa_pointers(fp_ctx *ctx, void **ptrs, uint32_t n_ptrs)
{
    FP_CHECK(ctx, ctx && ptrs, "null argument");
    if (int rc = ensure_current(ctx, 0, ctx->desc.capacity)) return rc;
    for (uint32_t f = 0; f < n_ptrs && f < fp::F_COUNT; ++f) ptrs[f] = ctx->s().u(f);
    return FP_OK;
}

// ---------------------------------------------------------------- met / pbl

static int upload_met_common(fp_ctx *ctx, const float *const f[6], hipMemcpyKind kind,
                             const float *level_heights)
{
    const uint64_t npts = (uint64_t)ctx->desc.met_nx * ctx->desc.met_ny * ctx->desc.met_nz;
    if (int rc = validate_levels(ctx, level_heights)) return rc;
    const float *src[6];
    if (kind == hipMemcpyHostToDevice) {
        if (!ctx->met_stage) FP_HIP(ctx, dalloc(&ctx->met_stage, npts * 6));
        for (int k = 0; k < 6; ++k) {
            FP_HIP(ctx, hipMemcpyAsync(ctx->met_stage + k * npts, f[k], npts * 4, kind, ctx->stream));
            src[k] = ctx->met_stage + k * npts;
        }
    } else {
        for (int k = 0; k < 6; ++k) src[k] = f[k];
    }
    FP_HIP(ctx, hipMemcpyAsync(ctx->level_heights, ctx->level_heights_host.data(),
                               ctx->level_heights_host.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    FP_HIP(ctx, fp::launch_pack_met(src[0], src[1], src[2], src[3], src[4], src[5], ctx->met,
                                    (uint32_t)npts, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->met_ready = true;
    return FP_OK;
}

int fp_met_upload_bracket(fp_ctx *ctx, const float *u0, const float *v0, const float *w0,
                          const float *u1, const float *v1, const float *w1,
                          const float *level_heights)
{
    FP_CHECK(ctx, ctx && u0 && v0 && w0 && u1 && v1 && w1 && level_heights, "null argument");
    const float *f[6] = {u0, v0, w0, u1, v1, w1};
    return upload_met_common(ctx, f, hipMemcpyHostToDevice, level_heights);
}

void *fp_host_alloc(uint64_t bytes)
{
    void *p = nullptr;
    if (bytes == 0 || hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

int fp_host_free(void *p)
{
    if (!p) return FP_OK;
    if (hipHostFree(p) != hipSuccess) return fail(nullptr, FP_E_DEVICE, "hipHostFree failed");
    return FP_OK;
}

int fp_met_prefetch_bracket(fp_ctx *ctx, const float *u0, const float *v0, const float *w0,
                            const float *u1, const float *v1, const float *w1,
                            const float *level_heights)
{
    FP_CHECK(ctx, ctx && u0 && v0 && w0 && u1 && v1 && w1 && level_heights, "null argument");
    if (ctx->prefetch_pending) return fail(ctx, FP_E_STATE, "a prefetched bracket is pending (fp_met_commit_prefetch)");
    if (int rc = check_levels(ctx, level_heights)) return rc;
    const uint64_t npts = (uint64_t)ctx->desc.met_nx * ctx->desc.met_ny * ctx->desc.met_nz;
    if (!ctx->copy_stream) {
        FP_HIP(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
        FP_HIP(ctx, hipEventCreateWithFlags(&ctx->prefetch_done, hipEventDisableTiming));
        FP_HIP(ctx, hipEventCreateWithFlags(&ctx->stage_free, hipEventDisableTiming));
    }
    if (!ctx->prefetch_stage) FP_HIP(ctx, dalloc(&ctx->prefetch_stage, npts * 6));
    // the previous commit's packing must have read the stage before it is overwritten
    if (ctx->stage_used) FP_HIP(ctx, hipStreamWaitEvent(ctx->copy_stream, ctx->stage_free, 0));
    const float *f[6] = {u0, v0, w0, u1, v1, w1};
    for (int k = 0; k < 6; ++k)
        FP_HIP(ctx, hipMemcpyAsync(ctx->prefetch_stage + k * npts, f[k], npts * 4, hipMemcpyHostToDevice,
                                   ctx->copy_stream));
    FP_HIP(ctx, hipEventRecord(ctx->prefetch_done, ctx->copy_stream));
    ctx->prefetch_lh.assign(level_heights, level_heights + ctx->desc.met_nz);
    ctx->prefetch_pending = true;
    return FP_OK;
}

int fp_met_commit_prefetch(fp_ctx *ctx)
{
    FP_CHECK(ctx, ctx, "null context");
    if (!ctx->prefetch_pending) return fail(ctx, FP_E_STATE, "no prefetched bracket (fp_met_prefetch_bracket)");
    const uint64_t npts = (uint64_t)ctx->desc.met_nx * ctx->desc.met_ny * ctx->desc.met_nz;
    FP_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->prefetch_done, 0));
    if (int rc = validate_levels(ctx, ctx->prefetch_lh.data())) return rc;
    FP_HIP(ctx, hipMemcpyAsync(ctx->level_heights, ctx->level_heights_host.data(),
                               ctx->level_heights_host.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    const float *s = ctx->prefetch_stage;
    FP_HIP(ctx, fp::launch_pack_met(s, s + npts, s + 2 * npts, s + 3 * npts, s + 4 * npts, s + 5 * npts, ctx->met,
                                    (uint32_t)npts, ctx->stream));
    FP_HIP(ctx, hipEventRecord(ctx->stage_free, ctx->stream));
    ctx->stage_used = true;
    ctx->prefetch_pending = false;
    ctx->met_ready = true;
    return FP_OK;
}

int fp_met_hybrid_to_height(fp_ctx *ctx, const float *ps, const float *temperature, const float *humidity,
                            const float *ak_pa, const float *bk, float *mean_heights, float *pressure_out,
                            float *height_out, fp_vt_error *err)
{
    FP_CHECK(ctx, ctx && ps && temperature && ak_pa && bk && mean_heights, "null argument");
    if (err) *err = fp_vt_error{FP_VT_OK, 0, 0, 0};
    const uint32_t ny = ctx->desc.met_ny, nz = ctx->desc.met_nz;
    const uint64_t ncol = (uint64_t)ctx->desc.met_nx * ny, ncell = ncol * nz;
    if (!ctx->vt_buf) {
        FP_HIP(ctx, dalloc(&ctx->vt_buf, ncol + 4 * ncell + 3 * (uint64_t)nz));
        FP_HIP(ctx, dalloc(&ctx->vt_err, 2));
    }
    float *d_ps = ctx->vt_buf, *d_t = d_ps + ncol, *d_q = d_t + ncell, *d_p = d_q + ncell, *d_z = d_p + ncell;
    float *d_ak = d_z + ncell, *d_bk = d_ak + nz, *d_mean = d_bk + nz;
    FP_HIP(ctx, hipMemcpyAsync(d_ps, ps, ncol * 4, hipMemcpyHostToDevice, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(d_t, temperature, ncell * 4, hipMemcpyHostToDevice, ctx->stream));
    if (humidity) FP_HIP(ctx, hipMemcpyAsync(d_q, humidity, ncell * 4, hipMemcpyHostToDevice, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(d_ak, ak_pa, nz * 4, hipMemcpyHostToDevice, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(d_bk, bk, nz * 4, hipMemcpyHostToDevice, ctx->stream));
    FP_HIP(ctx, hipMemsetAsync(ctx->vt_err, 0xff, 16, ctx->stream));
    FP_HIP(ctx, fp::launch_vertical_transform(d_ps, d_t, humidity ? d_q : nullptr, d_ak, d_bk, d_p, d_z, d_mean,
                                              ctx->vt_err, (uint32_t)ncol, nz, ctx->stream));
    unsigned long long keys[2];
    FP_HIP(ctx, hipMemcpyAsync(keys, ctx->vt_err, 16, hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(mean_heights, d_mean, nz * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (pressure_out) FP_HIP(ctx, hipMemcpyAsync(pressure_out, d_p, ncell * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (height_out) FP_HIP(ctx, hipMemcpyAsync(height_out, d_z, ncell * 4, hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const unsigned long long key = keys[0] != ~0ull ? keys[0] : keys[1];
    if (key != ~0ull) {
        const uint32_t c = (uint32_t)(key >> 32), k = (uint32_t)((key >> 8) & 0xffffffu), code = (uint32_t)(key & 0xffu);
        if (err) *err = fp_vt_error{code, c / ny, c % ny, k};
        static const char *what[5] = {"", "non-positive pressure", "pressure must decrease with level index",
                                      "invalid thermodynamic state", "geometric heights must increase with level index"};
        return fail(ctx, FP_E_INVALID, std::string(what[code < 5 ? code : 0]) + " at (i=" + std::to_string(c / ny) +
                                           ", j=" + std::to_string(c % ny) + ", k=" + std::to_string(k) + ")");
    }
    return FP_OK;
}

int fp_met_upload_bracket_dev(fp_ctx *ctx, const float *u0, const float *v0, const float *w0,
                              const float *u1, const float *v1, const float *w1,
                              const float *level_heights_host)
{
    FP_CHECK(ctx, ctx && u0 && v0 && w0 && u1 && v1 && w1 && level_heights_host, "null argument");
    const float *f[6] = {u0, v0, w0, u1, v1, w1};
    return upload_met_common(ctx, f, hipMemcpyDeviceToDevice, level_heights_host);
}

int fp_pbl_upload(fp_ctx *ctx, const float *ustar, const float *wstar, const float *hmix,
                  const float *oli, const float *precip)
{
    FP_CHECK(ctx, ctx && ustar && wstar && hmix && oli, "null argument");
    const uint64_t np = (uint64_t)ctx->desc.met_nx * ctx->desc.met_ny;
    if (!ctx->pbl_stage) FP_HIP(ctx, dalloc(&ctx->pbl_stage, np * 4));
    const float *f[4] = {ustar, wstar, hmix, oli};
    for (int k = 0; k < 4; ++k)
        FP_HIP(ctx, hipMemcpyAsync(ctx->pbl_stage + k * np, f[k], np * 4, hipMemcpyHostToDevice,
                                   ctx->stream));
    FP_HIP(ctx, fp::launch_pack_pbl(ctx->pbl_stage, ctx->pbl, (uint32_t)np, ctx->stream));
    if (precip) {
        FP_HIP(ctx, hipMemcpyAsync(ctx->precip, precip, np * 4, hipMemcpyHostToDevice, ctx->stream));
        ctx->precip_ready = true;
    }
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->pbl_ready = true;
    return FP_OK;
}

static int surface_upload(fp_ctx *ctx, int level, const float *const fields[FP_SURF_COUNT], int64_t time_s,
                          hipMemcpyKind kind)
{
    FP_CHECK(ctx, ctx && fields, "null argument");
    FP_CHECK(ctx, level == 0 || level == 1, "level must be 0 (t0) or 1 (t1)");
    for (int k = 0; k < FP_SURF_COUNT; ++k) FP_CHECK(ctx, fields[k], "null surface field");
    const uint64_t np = (uint64_t)ctx->desc.met_nx * ctx->desc.met_ny;
    if (!ctx->surf[level]) FP_HIP(ctx, dalloc(&ctx->surf[level], np * FP_SURF_COUNT));
    ctx->surf_ready[level] = false;
    for (int k = 0; k < FP_SURF_COUNT; ++k)
        FP_HIP(ctx, hipMemcpyAsync(ctx->surf[level] + k * np, fields[k], np * 4, kind, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->surf_time[level] = time_s;
    ctx->surf_ready[level] = true;
    return FP_OK;
}

int fp_surface_upload(fp_ctx *ctx, int level, const float *const fields[FP_SURF_COUNT], int64_t time_s)
{
    return surface_upload(ctx, level, fields, time_s, hipMemcpyHostToDevice);
}

int fp_surface_upload_dev(fp_ctx *ctx, int level, const float *const fields[FP_SURF_COUNT], int64_t time_s)
{
    return surface_upload(ctx, level, fields, time_s, hipMemcpyDeviceToDevice);
}

int fp_pbl_from_surface(fp_ctx *ctx, const fp_pbl_options *options, int64_t now_s, int strict,
                        float *passthrough_dev)
{
    FP_CHECK(ctx, ctx && options, "null argument");
    FP_CHECK(ctx, ctx->surf_ready[0] && ctx->surf_ready[1], "surface fields of both bracket ends are required");
    float alpha = 0.0f;
    if (fp_interpolation_alpha(ctx->surf_time[0], ctx->surf_time[1], now_s, strict, &alpha) != FP_OK)
        return fail(ctx, FP_E_INVALID, fp_last_error(nullptr));
    const uint64_t np = (uint64_t)ctx->desc.met_nx * ctx->desc.met_ny;
    const float *s0[FP_SURF_COUNT], *s1[FP_SURF_COUNT];
    for (int k = 0; k < FP_SURF_COUNT; ++k) {
        s0[k] = ctx->surf[0] + k * np;
        s1[k] = ctx->surf[1] + k * np;
    }
    ctx->pbl_ready = false;
    fp_ctx::Ev ev{};
    prof_begin(ctx, ev, FP_KERNEL_PBL);
    FP_HIP(ctx, fp::launch_pbl_surface(s0, s1, alpha, *options, ctx->pbl, passthrough_dev, (uint32_t)np, ctx->stream));
    prof_end(ctx, ev);
    ctx->pbl_ready = true;  // ordered before the next step on the context's stream
    return FP_OK;
}

int fp_pbl_download(fp_ctx *ctx, float *ustar, float *wstar, float *hmix, float *oli)
{
    FP_CHECK(ctx, ctx, "null context");
    FP_CHECK(ctx, ctx->pbl_ready, "no PBL state (fp_pbl_upload or fp_pbl_from_surface first)");
    const uint64_t np = (uint64_t)ctx->desc.met_nx * ctx->desc.met_ny;
    std::vector<float4> rec(2 * np);
    FP_HIP(ctx, hipMemcpyAsync(rec.data(), ctx->pbl, 2 * np * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (uint64_t c = 0; c < np; ++c) {
        const float4 r = rec[2 * c];
        if (ustar) ustar[c] = r.x;
        if (wstar) wstar[c] = r.y;
        if (hmix) hmix[c] = r.z;
        if (oli) oli[c] = r.w;
    }
    return FP_OK;
}

static int output_alloc(fp_ctx *ctx)
{
    const uint64_t ncell = (uint64_t)ctx->desc.out_nx * ctx->desc.out_ny * ctx->desc.out_nz;
    if (!ctx->out_acc_count) {
        FP_HIP(ctx, dalloc(&ctx->out_acc_count, ncell));
        FP_HIP(ctx, dalloc(&ctx->out_acc_mass, ncell));
        FP_HIP(ctx, hipMemsetAsync(ctx->out_acc_count, 0, ncell * 8, ctx->stream));
        FP_HIP(ctx, hipMemsetAsync(ctx->out_acc_mass, 0, ncell * 8, ctx->stream));
        ctx->out_samples = 0;
    }
    return FP_OK;
}

int fp_output_reset(fp_ctx *ctx)
{
    FP_CHECK(ctx, ctx && ctx->grid_counts, "context has no output grid");
    if (int rc = output_alloc(ctx)) return rc;
    const uint64_t ncell = (uint64_t)ctx->desc.out_nx * ctx->desc.out_ny * ctx->desc.out_nz;
    FP_HIP(ctx, hipMemsetAsync(ctx->out_acc_count, 0, ncell * 8, ctx->stream));
    FP_HIP(ctx, hipMemsetAsync(ctx->out_acc_mass, 0, ncell * 8, ctx->stream));
    ctx->out_samples = 0;
    return FP_OK;
}

int fp_output_sample(fp_ctx *ctx)
{
    FP_CHECK(ctx, ctx && ctx->grid_counts, "context has no output grid");
    FP_CHECK(ctx, ctx->out_samples < 0xffffffffu, "too many samples in one output interval");
    if (int rc = output_alloc(ctx)) return rc;
    const uint64_t ncell = (uint64_t)ctx->desc.out_nx * ctx->desc.out_ny * ctx->desc.out_nz;
    fp_ctx::Ev ev{};
    prof_begin(ctx, ev, FP_KERNEL_OUTPUT);
    FP_HIP(ctx, fp::launch_output_sample(ctx->grid_counts, ctx->grid_mass, ctx->out_acc_count, ctx->out_acc_mass,
                                         ncell, ctx->stream));
    prof_end(ctx, ev);
    ++ctx->out_samples;
    return FP_OK;
}

int fp_output_finalize(fp_ctx *ctx, const fp_output_geometry *geometry, uint32_t *count_out, float *mass_out,
                       float *conc_out, uint32_t *n_samples)
{
    FP_CHECK(ctx, ctx && ctx->grid_counts, "context has no output grid");
    FP_CHECK(ctx, ctx->out_samples > 0, "no samples in the output interval");
    FP_CHECK(ctx, count_out && mass_out, "null argument");
    const uint32_t nx = ctx->desc.out_nx, ny = ctx->desc.out_ny, nz = ctx->desc.out_nz;
    const uint64_t ncell = (uint64_t)nx * ny * nz;
    const float *oh = ctx->grid_params.outheights;
    if (conc_out) {
        FP_CHECK(ctx, geometry, "concentrations need the output grid geometry");
        FP_CHECK(ctx, std::isfinite(geometry->dx) && std::isfinite(geometry->dy) && geometry->dx > 0.0 &&
                          geometry->dy > 0.0 && std::isfinite(geometry->ylat0),
                 "output cell size must be positive and finite");
        FP_CHECK(ctx, oh[nz - 1] > 0.0f, "concentrations need metre output heights");
        for (uint32_t k = 0; k < nz; ++k)
            FP_CHECK(ctx, oh[k] > (k ? oh[k - 1] : 0.0f), "output heights must increase from 0");
    }
    if (!ctx->out_count) {
        FP_HIP(ctx, dalloc(&ctx->out_count, ncell));
        FP_HIP(ctx, dalloc(&ctx->out_mass, ncell));
    }
    if (conc_out) {
        if (!ctx->out_conc) {
            FP_HIP(ctx, dalloc(&ctx->out_conc, ncell));
            FP_HIP(ctx, dalloc(&ctx->out_vol, (uint64_t)ny * nz));
        }
        std::vector<float> vol((uint64_t)ny * nz);
        fp::output_volumes(*geometry, oh, ny, nz, vol.data());
        FP_HIP(ctx, hipMemcpyAsync(ctx->out_vol, vol.data(), vol.size() * 4, hipMemcpyHostToDevice, ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    FP_HIP(ctx, fp::launch_output_finalize(ctx->out_acc_count, ctx->out_acc_mass, ctx->out_samples,
                                           ctx->grid_params.mass_scale, ctx->out_vol, ny, nz, ctx->out_count,
                                           ctx->out_mass, conc_out ? ctx->out_conc : nullptr, ncell, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(count_out, ctx->out_count, ncell * 4, hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipMemcpyAsync(mass_out, ctx->out_mass, ncell * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (conc_out)
        FP_HIP(ctx, hipMemcpyAsync(conc_out, ctx->out_conc, ncell * 4, hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (n_samples) *n_samples = ctx->out_samples;
    return FP_OK;
}

int fp_write_netcdf(const char *path, const fp_netcdf_snapshot *s, int format)
{
    if (!path || !s) return fail(nullptr, FP_E_INVALID, "null argument");
    if (format != 2 && format != 5) return fail(nullptr, FP_E_INVALID, "format must be 2 (CDF-2) or 5 (CDF-5)");
    if (s->nx == 0 || s->ny == 0 || s->nz == 0) return fail(nullptr, FP_E_INVALID, "grid shape must be non-zero");
    if (!s->count || !s->mass_kg) return fail(nullptr, FP_E_INVALID, "count and mass_kg are required");
    if ((uint64_t)s->nx * s->ny * s->nz * 4 > 0xfffffffcull && format == 2)
        return fail(nullptr, FP_E_INVALID, "variable too large for CDF-2");
    std::string err;
    const int rc = fp::write_netcdf(path, *s, format, err);
    if (rc != FP_OK) return fail(nullptr, rc, err);
    return FP_OK;
}

int fp_forcing_upload(fp_ctx *ctx, const float *dry_vd, const float *wet_lambda,
                      const float *wet_fraction)
{
    FP_CHECK(ctx, ctx, "null context");
    const uint64_t cap = ctx->desc.capacity;
    struct { const float *src; float **dst; } items[3] = {
        {dry_vd, &ctx->dry_vd}, {wet_lambda, &ctx->wet_lambda}, {wet_fraction, &ctx->wet_fraction}};
    for (auto &it : items) {
        if (!it.src) continue;
        if (!*it.dst) FP_HIP(ctx, dalloc(it.dst, cap));
        FP_HIP(ctx, hipMemcpyAsync(*it.dst, it.src, cap * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

// ---------------------------------------------------------------- grid

static fp::GridArgs grid_args(const fp_ctx *ctx)
{
    return fp::make_grid_args(ctx->grid_counts, ctx->grid_mass, ctx->desc.out_nx, ctx->desc.out_ny,
                              ctx->desc.out_nz, ctx->grid_params, ctx->desc.met_nx, ctx->desc.met_ny);
}

int fp_grid_configure(fp_ctx *ctx, const fp_grid_params *params)
{
    FP_CHECK(ctx, ctx && params, "null argument");
    FP_CHECK(ctx, ctx->grid_counts, "context has no output grid");
    FP_CHECK(ctx, params->refine_x >= 1 && params->refine_y >= 1, "refine factors must be >= 1");
    FP_CHECK(ctx, params->species < ctx->desc.n_species, "species index beyond n_species");
    FP_CHECK(ctx, std::isfinite(params->mass_scale), "mass_scale must be finite");
    ctx->grid_params = *params;
    ctx->grid_configured = true;
    return FP_OK;
}

int fp_grid_clear(fp_ctx *ctx)
{
    FP_CHECK(ctx, ctx && ctx->grid_counts, "context has no output grid");
    const uint64_t ncell = (uint64_t)ctx->desc.out_nx * ctx->desc.out_ny * ctx->desc.out_nz;
    FP_HIP(ctx, hipMemsetAsync(ctx->grid_counts, 0, ncell * 4, ctx->stream));
    FP_HIP(ctx, hipMemsetAsync(ctx->grid_mass, 0, ncell * 4, ctx->stream));
    return FP_OK;
}

int fp_grid_accumulate(fp_ctx *ctx)
{
    FP_CHECK(ctx, ctx && ctx->grid_counts, "context has no output grid");
    fp_ctx::Ev ev{};
    prof_begin(ctx, ev, FP_KERNEL_GRID);
    FP_HIP(ctx, fp::launch_grid(ctx->s(), (uint32_t)ctx->dispatch, grid_args(ctx), ctx->stream));
    prof_end(ctx, ev);
    return FP_OK;
}

int fp_grid_download(fp_ctx *ctx, uint32_t *counts, uint32_t *mass_fp)
{
    FP_CHECK(ctx, ctx && ctx->grid_counts, "context has no output grid");
    const uint64_t ncell = (uint64_t)ctx->desc.out_nx * ctx->desc.out_ny * ctx->desc.out_nz;
    if (counts)
        FP_HIP(ctx, hipMemcpyAsync(counts, ctx->grid_counts, ncell * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (mass_fp)
        FP_HIP(ctx, hipMemcpyAsync(mass_fp, ctx->grid_mass, ncell * 4, hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FP_OK;
}

// ---------------------------------------------------------------- step

}  // extern "C"

// Validated step arguments of fp_step (and fp_step_vjp) on the context's
// current particle set.
static int build_step_args(fp_ctx *ctx, const fp_step_params *p, fp::StepArgs &a)
{
    FP_CHECK(ctx, ctx && p, "null argument");
    if (!ctx->met_ready) return fail(ctx, FP_E_STATE, "fp_step before fp_met_upload_bracket");
    if (!ctx->pbl_ready) return fail(ctx, FP_E_STATE, "fp_step before fp_pbl_upload");
    FP_CHECK(ctx, std::isfinite(p->dt_seconds) && p->dt_seconds > 0.0f, "invalid dt_seconds");
    FP_CHECK(ctx, p->n_substeps <= 4, "n_substeps must be 0..4");
    FP_CHECK(ctx, p->direction == 1 || p->direction == -1, "direction must be +1 or -1");
    FP_CHECK(ctx, p->dry_mode <= FP_DRY_PER_SLOT, "unknown dry_mode");
    FP_CHECK(ctx, p->wet_mode <= FP_WET_PRECIP_FIELD, "unknown wet_mode");
    FP_CHECK(ctx, p->loss_mode <= FP_LOSS_MIN_MASS, "unknown loss_mode");
    if (p->dry_mode == FP_DRY_PER_SLOT && !ctx->dry_vd)
        return fail(ctx, FP_E_STATE, "per-slot dry deposition without fp_forcing_upload");
    if (p->wet_mode == FP_WET_PER_SLOT && !(ctx->wet_lambda && ctx->wet_fraction))
        return fail(ctx, FP_E_STATE, "per-slot wet deposition without fp_forcing_upload");
    if (p->wet_mode == FP_WET_PRECIP_FIELD && !ctx->precip_ready)
        return fail(ctx, FP_E_STATE, "precip-field wet deposition without a precip upload");
    if (p->grid_fused && !ctx->grid_configured)
        return fail(ctx, FP_E_STATE, "fused gridding before fp_grid_configure");

    a = fp::StepArgs{};
    a.s = ctx->s();
    a.n = (uint32_t)ctx->dispatch;
    a.n_species = ctx->desc.n_species;
    a.met = ctx->met;
    a.nx = ctx->desc.met_nx;
    a.ny = ctx->desc.met_ny;
    a.nz = ctx->desc.met_nz;
    a.level_heights = ctx->level_heights;
    a.has_heights = ctx->has_heights;
    a.zmax = ctx->zmax;
    a.pbl = ctx->pbl;
    a.precip = ctx->precip;
    a.dry_vd_slots = ctx->dry_vd;
    a.wet_lambda_slots = ctx->wet_lambda;
    a.wet_fraction_slots = ctx->wet_fraction;
    fp::fill_step_scalars(a, *p);
    if (a.grid_fused) a.grid = grid_args(ctx);
    a.ppt = ctx->ppt;
    return FP_OK;
}

extern "C" {

int fp_step(fp_ctx *ctx, const fp_step_params *p, uint32_t next_counter[4])
{
    fp::StepArgs a;
    if (const int rc = build_step_args(ctx, p, a)) return rc;

    fp_ctx::Ev ev{};
    prof_begin(ctx, ev, FP_KERNEL_STEP);
    FP_HIP(ctx, fp::launch_step(a, ctx->block_size, ctx->stream));
    prof_end(ctx, ev);

    if (next_counter)
        fp::counter_advance(p->counter, fp::counter_per_step(ctx->dispatch, p->n_substeps), next_counter);
    return FP_OK;
}

int fp_step_vjp(fp_ctx *ctx, const fp_step_params *p, const float *lam_out, float *lam_in, double *wind_u,
                double *wind_v)
{
    fp::StepArgs a;
    if (const int rc = build_step_args(ctx, p, a)) return rc;
    FP_CHECK(ctx, lam_out && lam_in, "null argument");
    FP_CHECK(ctx, (wind_u == nullptr) == (wind_v == nullptr), "wind_u and wind_v go together");
    a.grid_fused = 0;
    const uint64_t n = ctx->dispatch;
    if (n == 0) return FP_OK;
    const uint64_t nxy = (uint64_t)ctx->desc.met_nx * ctx->desc.met_ny;
    float *lam = nullptr;
    double *au = nullptr, *av = nullptr;
    FP_HIP(ctx, dalloc(&lam, n * 7u));
    auto cleanup = [&]() {
        if (lam) (void)hipFree(lam);
        if (au) (void)hipFree(au);
        if (av) (void)hipFree(av);
    };
    hipError_t h = hipMemcpyAsync(lam, lam_out, n * 7u * 4u, hipMemcpyHostToDevice, ctx->stream);
    if (h == hipSuccess && wind_u) {
        h = dalloc(&au, nxy);
        if (h == hipSuccess) h = dalloc(&av, nxy);
        if (h == hipSuccess) h = hipMemsetAsync(au, 0, nxy * 8u, ctx->stream);
        if (h == hipSuccess) h = hipMemsetAsync(av, 0, nxy * 8u, ctx->stream);
    }
    if (h == hipSuccess) h = fp::launch_adj_step(a, lam, n, au, av, ctx->stream);
    if (h == hipSuccess) h = hipMemcpyAsync(lam_in, lam, n * 7u * 4u, hipMemcpyDeviceToHost, ctx->stream);
    if (h == hipSuccess && wind_u) h = hipMemcpyAsync(wind_u, au, nxy * 8u, hipMemcpyDeviceToHost, ctx->stream);
    if (h == hipSuccess && wind_u) h = hipMemcpyAsync(wind_v, av, nxy * 8u, hipMemcpyDeviceToHost, ctx->stream);
    if (h == hipSuccess) h = hipStreamSynchronize(ctx->stream);
    cleanup();
    FP_HIP(ctx, h);
    return FP_OK;
}


// ---------------------------------------------------------------- compaction / sort

int fp_compact(fp_ctx *ctx, uint64_t *active_count, uint32_t *survivor_map)
{
    FP_CHECK(ctx, ctx, "null context");
    const uint32_t n = (uint32_t)ctx->dispatch;
    const int other = 1 - ctx->cur;
    fp_ctx::Ev ev{};
    prof_begin(ctx, ev, FP_KERNEL_COMPACT);
    FP_HIP(ctx, fp::launch_compact(ctx->soa[ctx->cur], ctx->soa[other], n, ctx->block_counts,
                                   ctx->survivor_map, ctx->active_count_dev, ctx->stream));
    prof_end(ctx, ev);
    FP_HIP(ctx, hipMemcpyAsync(ctx->active_count_host, ctx->active_count_dev, 4,
                               hipMemcpyDeviceToHost, ctx->stream));
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t a = *ctx->active_count_host;
    ctx->segs.assign(0, n, other);
    ctx->cur = other;
    ctx->dispatch = a;  // ParticleStore::reset_after_compaction (particles/mod.rs:600-605)
    if (survivor_map && a)
        FP_HIP(ctx, hipMemcpy(survivor_map, ctx->survivor_map, (uint64_t)a * 4, hipMemcpyDeviceToHost));
    if (active_count) *active_count = a;
    return FP_OK;
}

int fp_sort_morton(fp_ctx *ctx, const fp_sort_params *p, uint32_t *new_to_old)
{
    FP_CHECK(ctx, ctx && p, "null argument");
    FP_CHECK(ctx, p->bits_per_axis >= 1 && p->bits_per_axis <= 21, "bits_per_axis must be 1..21");
    const uint32_t n = (uint32_t)ctx->dispatch;
    if (n == 0) return FP_OK;
    if (int rc = ensure_stage(ctx, n)) return rc;
    fp_ctx::Ev ev{};
    prof_begin(ctx, ev, FP_KERNEL_SORT);
    FP_HIP(ctx, fp::sort_morton_range(ctx->sortw, *p, ctx->soa[ctx->cur], ctx->soa[1 - ctx->cur], ctx->aos_stage, n,
                                      ctx->stream));
    prof_end(ctx, ev);
    const int other = 1 - ctx->cur;
    ctx->segs.assign(0, n, other);
    ctx->cur = other;
    if (new_to_old) {
        FP_HIP(ctx, hipMemcpyAsync(new_to_old, ctx->sortw.vals_alt, (uint64_t)n * 4,
                                   hipMemcpyDeviceToHost, ctx->stream));
        FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return FP_OK;
}

// ---------------------------------------------------------------- philox

int fp_philox_blocks(fp_ctx *ctx, const uint32_t counter[4], const uint32_t key[2], uint64_t n,
                     uint32_t *out)
{
    FP_CHECK(ctx, ctx && counter && key && (out || n == 0), "null argument");
    if (n == 0) return FP_OK;
    uint32_t *d = nullptr;
    FP_HIP(ctx, dalloc(&d, n * 4));
    hipError_t e = fp::launch_philox_blocks(counter, key, n, d, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, n * 16, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    FP_HIP(ctx, e);
    return FP_OK;
}

int fp_release_inject(fp_ctx *ctx, const fp_grid_domain *grid, const fp_release *rel, uint32_t n_rel,
                      uint64_t *released, int64_t now_s, const uint32_t *slots, uint64_t n_slots,
                      uint64_t first_slot, uint64_t *injected)
{
    FP_CHECK(ctx, ctx && grid && (rel || n_rel == 0) && (released || n_rel == 0), "null argument");
    if (injected) *injected = 0;
    FP_CHECK(ctx, grid->nx >= 1 && grid->ny >= 1 && grid->dx != 0.0 && grid->dy != 0.0, "invalid grid domain");
    FP_CHECK(ctx, now_s >= INT32_MIN && now_s <= INT32_MAX, "timestamp cannot be represented as i32 seconds");
    for (uint32_t k = 0; k < n_rel; ++k) {
        FP_CHECK(ctx, rel[k].particle_count >= 1, "release particle_count must be >= 1");
        FP_CHECK(ctx, rel[k].lon_min <= rel[k].lon_max, "invalid bounds: longitude min > max");
        FP_CHECK(ctx, rel[k].lat_min <= rel[k].lat_max, "invalid bounds: latitude min > max");
        FP_CHECK(ctx, rel[k].z_min <= rel[k].z_max, "invalid bounds: height min > max");
    }
    const uint64_t cap = ctx->desc.capacity;
    uint64_t max_particles;
    const uint32_t *slots_dev = nullptr;
    if (slots) {
        for (uint64_t k = 0; k < n_slots; ++k) FP_CHECK(ctx, slots[k] < cap, "slot index out of range");
        max_particles = n_slots;
        if (n_slots) {
            uint64_t lo = UINT64_MAX, hi = 0;
            for (uint64_t k = 0; k < n_slots; ++k) {
                lo = std::min<uint64_t>(lo, slots[k]);
                hi = std::max<uint64_t>(hi, slots[k]);
            }
            if (int rc = ensure_current(ctx, lo, hi + 1)) return rc;
            if (int rc = ensure_slot_stage(ctx, n_slots)) return rc;
            FP_HIP(ctx, hipMemcpyAsync(ctx->slot_stage, slots, n_slots * 4, hipMemcpyHostToDevice, ctx->stream));
            slots_dev = ctx->slot_stage;
        }
    } else {
        FP_CHECK(ctx, first_slot <= cap, "first_slot beyond capacity");
        max_particles = cap - first_slot;
    }
    uint64_t n = 0;
    int status = 0;
    FP_HIP(ctx, fp::release_plan_and_launch(*grid, rel, n_rel, released, now_s, (int32_t)now_s, slots_dev,
                                            first_slot, max_particles, ctx->s(), &ctx->release_jobs,
                                            &ctx->release_jobs_bytes, ctx->stream, &n, &status));
    // every word of the written slots is new: the current set owns them now
    if (!slots) ctx->segs.assign(first_slot, first_slot + n, ctx->cur);
    if (injected) *injected = n;
    if (status == 1) return fail(ctx, FP_E_INVALID, "release sampled outside the grid domain");
    if (status == 2) return fail(ctx, FP_E_INVALID, "particle store full");
    return FP_OK;
}

int fp_interpolation_alpha(int64_t t0_s, int64_t t1_s, int64_t target_s, int strict, float *alpha)
{
    if (!alpha) return fail(nullptr, FP_E_INVALID, "null argument");
    if (t1_s <= t0_s) return fail(nullptr, FP_E_INVALID, "invalid time bracket: t1 <= t0");
    int64_t t = target_s;
    if (t < t0_s || t > t1_s) {
        if (strict) return fail(nullptr, FP_E_INVALID, "target time outside the met bracket");
        t = std::min(std::max(t, t0_s), t1_s);
    }
    *alpha = (float)((double)(t - t0_s) / (double)(t1_s - t0_s));
    return FP_OK;
}

int fp_math_probe(fp_ctx *ctx, int fn, const float *x, const float *y, float *out, uint64_t n)
{
    FP_CHECK(ctx, ctx && x && y && (out || n == 0), "null argument");
    FP_CHECK(ctx, fn >= FP_PROBE_EXP && fn <= FP_PROBE_LEVEL, "unknown probe function");
    if (fn == FP_PROBE_LEVEL && !ctx->met_ready) return fail(ctx, FP_E_STATE, "no met bracket uploaded");
    if (n == 0) return FP_OK;
    float *d = nullptr;
    FP_HIP(ctx, dalloc(&d, n * 3));
    hipError_t e = hipMemcpyAsync(d, x, n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d + n, y, n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = fn == FP_PROBE_LEVEL
                ? fp::launch_level_probe(ctx->level_heights, ctx->desc.met_nz, ctx->has_heights, d, d + 2 * n, n,
                                         ctx->stream)
                : fp::launch_math_probe(fn, d, d + n, d + 2 * n, n, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d + 2 * n, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    FP_HIP(ctx, e);
    return FP_OK;
}

int fp_math_sweep(fp_ctx *ctx, int fn, uint32_t first_bits, uint32_t last_bits, uint64_t *failures,
                  uint32_t *first_failure)
{
    FP_CHECK(ctx, ctx && failures && first_failure, "null argument");
    FP_CHECK(ctx, fn == FP_SWEEP_SQRT_STEP, "unknown sweep function");
    FP_CHECK(ctx, first_bits <= last_bits, "empty bit-pattern range");
    uint32_t *d = nullptr;
    FP_HIP(ctx, dalloc(&d, 4));
    auto *fails = reinterpret_cast<unsigned long long *>(d);
    hipError_t e = fp::launch_math_sweep(fn, first_bits, last_bits, fails, d + 2, ctx->stream);
    unsigned long long hf = 0;
    uint32_t hw = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&hf, fails, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&hw, d + 2, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    FP_HIP(ctx, e);
    *failures = hf;
    *first_failure = hw;
    return FP_OK;
}

// ---------------------------------------------------------------- timing / tuning

int fp_profile_enable(fp_ctx *ctx, int enable)
{
    FP_CHECK(ctx, ctx, "null context");
    ctx->profiling = enable != 0;
    return FP_OK;
}

static int collect_events(fp_ctx *ctx)
{
    FP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &ev : ctx->events) {
        float ms = 0.0f;
        FP_HIP(ctx, hipEventElapsedTime(&ms, ev.a, ev.b));
        ctx->prof_ms[ev.kernel] += ms;
        ctx->prof_n[ev.kernel] += 1;
        ctx->free_events.push_back(ev);
    }
    ctx->events.clear();
    return FP_OK;
}

int fp_profile_read(fp_ctx *ctx, int kernel, double *total_ms, uint64_t *launches)
{
    FP_CHECK(ctx, ctx, "null context");
    FP_CHECK(ctx, kernel >= 0 && kernel < FP_KERNEL_COUNT, "unknown kernel id");
    if (int rc = collect_events(ctx)) return rc;
    if (total_ms) *total_ms = ctx->prof_ms[kernel];
    if (launches) *launches = ctx->prof_n[kernel];
    return FP_OK;
}

int fp_profile_reset(fp_ctx *ctx)
{
    FP_CHECK(ctx, ctx, "null context");
    if (int rc = collect_events(ctx)) return rc;
    for (int k = 0; k < FP_KERNEL_COUNT; ++k) { ctx->prof_ms[k] = 0.0; ctx->prof_n[k] = 0; }
    return FP_OK;
}

int fp_set_launch(fp_ctx *ctx, uint32_t block_size, uint32_t particles_per_thread)
{
    FP_CHECK(ctx, ctx, "null context");
    FP_CHECK(ctx, block_size == 64 || block_size == 128 || block_size == 256 || block_size == 512 ||
                      block_size == 1024,
             "block_size must be 64, 128, 256, 512 or 1024");
    FP_CHECK(ctx, particles_per_thread >= 1 && particles_per_thread <= 16,
             "particles_per_thread must be 1..16");
    ctx->block_size = block_size;
    ctx->ppt = particles_per_thread;
    return FP_OK;
}

}  // extern "C"

namespace fp {
void set_context_free_error(const char *msg) { set_create_error(msg); }

uint32_t level_table(const float *lh, uint32_t nz, std::vector<float> &out)
{
    out.assign(lh, lh + nz);
    out.resize(level_table_words(nz), 0.0f);
    out[nz] = out[nz + 1] = std::numeric_limits<float>::infinity();
    const float top = lh[nz - 1];
    const float inv_w = top > 0.0f ? (float)kLevelBuckets / top : 0.0f;
    out[nz + 2] = inv_w;
    uint16_t first[kLevelBuckets];
    uint32_t last_first = 0;
    for (uint32_t b = 0; b < kLevelBuckets; ++b) {
        // the device maps z to band b = (uint)(z * inv_w) with one rounding of
        // the product, so every z in band b satisfies z >= b / inv_w (1 - 2^-24)
        const double zlow = inv_w > 0.0f ? (double)b / (double)inv_w * (1.0 - 1.0e-6) : 0.0;
        uint32_t k = 1;
        while (k < nz && (double)lh[k] < zlow) ++k;
        first[b] = (uint16_t)std::min(k, nz);
        last_first = std::max(last_first, (uint32_t)first[b]);
    }
    std::memcpy(out.data() + nz + 3, first, sizeof first);
    if (!(top > 0.0f)) return kLevelIndex;
    // the banded lookup (level_of_banded) needs the below-first-level quotient
    // z / h0 (a positive h0) and every band's first candidate below the top
    // level, so that it reads at most the two +inf words past the heights
    return nz >= 2 && lh[0] > 0.0f && last_first <= nz - 1 ? kLevelHeightsBanded : kLevelHeights;
}
}  // namespace fp
