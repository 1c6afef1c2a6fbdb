// fp_tune.hip - launch-shape and re-sort-interval autotuner (fp_autotune).
//
// Mirrors the reference's workgroup autotuner (gpu/workgroup.rs:133-147,
// 342-445): every candidate gets warm-up runs and measured runs, the median
// of the measured runs is compared, ties keep the earlier candidate, and the
// winner is cached per device in a JSON file.  Candidates here are the step
// kernel's block size and particles per thread, then - with the winning
// shape - the Morton re-sort interval, timed with amortised sort cost.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "fp_internal.hpp"

namespace {

struct Snapshot {
    uint32_t *soa = nullptr;
    uint32_t *grid = nullptr;
    uint64_t n = 0, ncell = 0;
};

hipError_t take(fp_ctx *ctx, Snapshot &s)
{
    s.n = ctx->dispatch;
    s.ncell = ctx->grid_counts ? (uint64_t)ctx->desc.out_nx * ctx->desc.out_ny * ctx->desc.out_nz : 0;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&s.soa), std::max<uint64_t>(s.n, 1) * fp::F_COUNT * 4);
    if (e != hipSuccess) return e;
    if (s.ncell) {
        e = hipMalloc(reinterpret_cast<void **>(&s.grid), s.ncell * 8);
        if (e != hipSuccess) return e;
        e = hipMemcpyAsync(s.grid, ctx->grid_counts, s.ncell * 4, hipMemcpyDeviceToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(s.grid + s.ncell, ctx->grid_mass, s.ncell * 4, hipMemcpyDeviceToDevice,
                               ctx->stream);
        if (e != hipSuccess) return e;
    }
    const fp::Soa &cur = ctx->soa[ctx->cur];
    for (uint32_t f = 0; f < fp::F_COUNT && s.n; ++f) {
        e = hipMemcpyAsync(s.soa + f * s.n, cur.u(f), s.n * 4, hipMemcpyDeviceToDevice, ctx->stream);
        if (e != hipSuccess) return e;
    }
    return hipStreamSynchronize(ctx->stream);
}

hipError_t restore(fp_ctx *ctx, const Snapshot &s)
{
    ctx->dispatch = s.n;
    const fp::Soa &cur = ctx->soa[ctx->cur];
    hipError_t e = hipSuccess;
    for (uint32_t f = 0; f < fp::F_COUNT && s.n && e == hipSuccess; ++f)
        e = hipMemcpyAsync(cur.u(f), s.soa + f * s.n, s.n * 4, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess && s.ncell) {
        e = hipMemcpyAsync(ctx->grid_counts, s.grid, s.ncell * 4, hipMemcpyDeviceToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ctx->grid_mass, s.grid + s.ncell, s.ncell * 4, hipMemcpyDeviceToDevice,
                               ctx->stream);
    }
    if (e == hipSuccess && s.n) ctx->segs.assign(0, s.n, ctx->cur);
    return e;
}

void release(Snapshot &s)
{
    if (s.soa) (void)hipFree(s.soa);
    if (s.grid) (void)hipFree(s.grid);
    s.soa = s.grid = nullptr;
}

std::string cache_key(const fp_ctx *ctx, const fp_step_params &p)
{
    hipDeviceProp_t prop{};
    (void)hipGetDeviceProperties(&prop, ctx->device);
    std::ostringstream k;
    k << prop.gcnArchName << "|cu" << prop.multiProcessorCount << "|n" << ctx->dispatch << "|met"
      << ctx->desc.met_nx << "x" << ctx->desc.met_ny << "x" << ctx->desc.met_nz << "|sp"
      << ctx->desc.n_species << "|sub" << p.n_substeps << "|dry" << p.dry_mode << "|wet" << p.wet_mode
      << "|grid" << p.grid_fused;
    return k.str();
}

// One entry per line: {"key": "...", "block": B, "ppt": P, "sort": S, "ms": T, "ms_sort": U}
bool cache_lookup(const char *path, const std::string &key, fp_autotune_result &r)
{
    if (!path) return false;
    std::ifstream in(path);
    std::string line;
    const std::string needle = "\"key\": \"" + key + "\"";
    while (std::getline(in, line)) {
        if (line.find(needle) == std::string::npos) continue;
        unsigned b = 0, p = 0, s = 0;
        double ms = 0, mss = 0;
        const size_t at = line.find("\"block\"");
        if (at == std::string::npos) continue;
        if (std::sscanf(line.c_str() + at, "\"block\": %u, \"ppt\": %u, \"sort\": %u, \"ms\": %lf, \"ms_sort\": %lf",
                        &b, &p, &s, &ms, &mss) == 5) {
            r.block_size = b;
            r.particles_per_thread = p;
            r.sort_interval = s;
            r.step_ms = ms;
            r.step_ms_with_sort = mss;
            r.from_cache = 1;
            return true;
        }
    }
    return false;
}

void cache_store(const char *path, const std::string &key, const fp_autotune_result &r)
{
    if (!path) return;
    std::vector<std::string> keep;
    {
        std::ifstream in(path);
        std::string line;
        const std::string needle = "\"key\": \"" + key + "\"";
        while (std::getline(in, line))
            if (line.find(needle) == std::string::npos && !line.empty()) keep.push_back(line);
    }
    std::ofstream out(path, std::ios::trunc);
    for (const auto &l : keep) out << l << "\n";
    char buf[512];
    std::snprintf(buf, sizeof buf,
                  "{\"key\": \"%s\", \"block\": %u, \"ppt\": %u, \"sort\": %u, \"ms\": %.6f, \"ms_sort\": %.6f}",
                  key.c_str(), r.block_size, r.particles_per_thread, r.sort_interval, r.step_ms,
                  r.step_ms_with_sort);
    out << buf << "\n";
}

}  // namespace

extern "C" int fp_autotune(fp_ctx *ctx, const fp_step_params *base, const fp_autotune_params *params,
                           fp_autotune_result *out)
{
    if (!ctx || !base || !params || !out) return FP_E_INVALID;
    if (params->n_sort_intervals > 8 || params->measured_steps == 0) {
        ctx->err = "fp_autotune: measured_steps must be >= 1 and n_sort_intervals <= 8";
        return FP_E_INVALID;
    }
    std::memset(out, 0, sizeof *out);
    const std::string key = cache_key(ctx, *base);
    if (cache_lookup(params->cache_path, key, *out)) return fp_set_launch(ctx, out->block_size, out->particles_per_thread);

    Snapshot snap;
    hipError_t he = take(ctx, snap);
    if (he != hipSuccess) {
        release(snap);
        ctx->err = std::string("fp_autotune snapshot: ") + hipGetErrorString(he);
        return FP_E_NOMEM;
    }
    hipEvent_t ea, eb;
    (void)hipEventCreate(&ea);
    (void)hipEventCreate(&eb);
    const bool was_profiling = ctx->profiling;
    ctx->profiling = false;
    int rc = FP_OK;

    auto run_steps = [&](uint32_t steps, uint32_t sort_every, std::vector<double> *per_step,
                         double *total_ms) -> int {
        fp_step_params p = *base;
        for (uint32_t k = 0; k < steps; ++k) {
            if (sort_every && k % sort_every == 0) {
                fp_sort_params sp{};
                sp.bits_per_axis = params->sort_bits ? params->sort_bits : 10;
                if (int r = fp_sort_morton(ctx, &sp, nullptr)) return r;
            }
            if (per_step) (void)hipEventRecord(ea, ctx->stream);
            uint32_t nxt[4];
            if (int r = fp_step(ctx, &p, nxt)) return r;
            std::memcpy(p.counter, nxt, sizeof nxt);
            if (per_step) {
                (void)hipEventRecord(eb, ctx->stream);
                (void)hipEventSynchronize(eb);
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, ea, eb);
                per_step->push_back(ms);
            }
        }
        if (total_ms) *total_ms = 0.0;
        return FP_OK;
    };

    // 1) launch shape, no re-sort
    static const uint32_t blocks[] = {64, 128, 256, 512};
    static const uint32_t ppts[] = {1, 2, 4};
    double best = 1e300;
    uint32_t best_b = ctx->block_size, best_p = ctx->ppt;
    for (uint32_t b : blocks) {
        for (uint32_t pp : ppts) {
            if ((rc = fp_set_launch(ctx, b, pp))) goto done;
            if ((he = restore(ctx, snap)) != hipSuccess) { rc = FP_E_DEVICE; goto done; }
            if ((rc = run_steps(params->warmup_steps, 0, nullptr, nullptr))) goto done;
            std::vector<double> t;
            if ((rc = run_steps(params->measured_steps, 0, &t, nullptr))) goto done;
            std::sort(t.begin(), t.end());
            const double med = t[t.size() / 2];
            if (med < best) { best = med; best_b = b; best_p = pp; }
        }
    }
    if ((rc = fp_set_launch(ctx, best_b, best_p))) goto done;
    out->block_size = best_b;
    out->particles_per_thread = best_p;
    out->step_ms = best;
    out->step_ms_with_sort = best;

    // 2) re-sort interval with the chosen shape (amortised sort cost included)
    if (params->n_sort_intervals) {
        const uint32_t steps = params->sort_trial_steps ? params->sort_trial_steps : 12;
        double best_s = 1e300;
        for (uint32_t c = 0; c < params->n_sort_intervals; ++c) {
            const uint32_t s = params->sort_intervals[c];
            if ((he = restore(ctx, snap)) != hipSuccess) { rc = FP_E_DEVICE; goto done; }
            if ((rc = run_steps(params->warmup_steps, 0, nullptr, nullptr))) goto done;
            (void)hipEventRecord(ea, ctx->stream);
            if ((rc = run_steps(steps, s, nullptr, nullptr))) goto done;
            (void)hipEventRecord(eb, ctx->stream);
            (void)hipEventSynchronize(eb);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ea, eb);
            const double per = ms / steps;
            if (per < best_s) { best_s = per; out->sort_interval = s; }
        }
        out->step_ms_with_sort = best_s;
    }
    cache_store(params->cache_path, key, *out);

done:
    if (restore(ctx, snap) != hipSuccess && rc == FP_OK) rc = FP_E_DEVICE;
    (void)hipStreamSynchronize(ctx->stream);
    release(snap);
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    ctx->profiling = was_profiling;
    if (rc != FP_OK && ctx->err.empty()) ctx->err = "fp_autotune failed";
    return rc;
}
