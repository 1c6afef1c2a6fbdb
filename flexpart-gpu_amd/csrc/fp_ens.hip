// fp_ens.hip - batched ensembles: many independent members, one launch per
// sync step, optionally replayed from a captured hipGraph (fp_ens_* in
// include/flexpart_hip.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fp_internal.hpp"
#include "fp_math.h"

static_assert(sizeof(fp_ens_desc) == 40, "fp_ens_desc layout is part of the ABI");

struct fp_ens {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    uint32_t n_members = 0;
    uint32_t nx = 0, ny = 0, nz = 0, n_species = 1;
    uint32_t ox = 0, oy = 0, oz = 0;
    std::vector<uint64_t> cap, off;
    uint64_t total = 0, max_cap = 0;
    fp::Soa soa;

    float4 *met = nullptr;  // n_members * npts * 2
    float *met_stage = nullptr;  // 6 * npts floats for host uploads
    float *level_heights = nullptr;
    std::vector<float> lh_host;
    uint32_t has_heights = fp::kLevelIndex;
    float zmax = 0.0f;
    std::vector<char> met_ready;

    float4 *pbl = nullptr;
    float *precip = nullptr;
    float *pbl_stage = nullptr;
    bool pbl_ready = false, precip_ready = false;

    uint32_t *grid_counts = nullptr, *grid_mass = nullptr;
    fp_grid_params gp{};
    bool grid_configured = false;

    fp_particle *aos_stage = nullptr;

    fp::EnsMember *d_members = nullptr;
    fp::EnsBlock *d_blocks = nullptr;
    uint64_t d_blocks_cap = 0;
    fp::EnsClock *d_clock = nullptr;
    std::vector<fp::EnsMember> members;
    uint32_t nblocks = 0;
    bool configured = false, sub4 = false;
    uint64_t steps_done = 0;

    uint32_t block = 256, ppt = 1;
    fp::SortWork sortw;  // allocated by the first fp_ens_sort_morton
    // captured step sequences: [0] one step, [1] `multi_steps` consecutive
    // steps (an output interval replayed as one graph launch)
    hipGraph_t graph[2] = {};
    hipGraphExec_t exec[2] = {};
    uint32_t multi_steps = 0;
    // the launch a graph was captured with: table addresses and launch
    // shape (the tables' contents are read at run time, so a re-configured
    // ensemble with the same signature replays its graphs unchanged)
    struct LaunchSig {
        const void *members, *blocks, *clock;
        uint32_t nblocks, block, ppt;
        bool sub4;
        bool operator==(const LaunchSig &o) const
        {
            return members == o.members && blocks == o.blocks && clock == o.clock && nblocks == o.nblocks &&
                   block == o.block && ppt == o.ppt && sub4 == o.sub4;
        }
    } graph_sig[2] = {};
    LaunchSig sig() const { return LaunchSig{d_members, d_blocks, d_clock, nblocks, block, ppt, sub4}; }
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    double last_ms = 0.0;

    uint64_t npts() const { return (uint64_t)nx * ny * nz; }
    uint64_t ncell() const { return (uint64_t)ox * oy * oz; }
};

namespace {

thread_local std::string g_ens_error;

int efail(fp_ens *e, int code, const std::string &msg)
{
    if (e) e->err = msg; else g_ens_error = msg;
    return code;
}

#define E_HIP(e, expr)                                                                        \
    do {                                                                                      \
        hipError_t h_ = (expr);                                                               \
        if (h_ != hipSuccess)                                                                 \
            return efail((e), h_ == hipErrorOutOfMemory ? FP_E_NOMEM : FP_E_DEVICE,           \
                         std::string(#expr) + ": " + hipGetErrorString(h_));                  \
    } while (0)
#define E_CHECK(e, cond, msg)                                                                 \
    do {                                                                                      \
        if (!(cond)) return efail((e), FP_E_INVALID, (msg));                                  \
    } while (0)
#define E_STATE(e, cond, msg)                                                                 \
    do {                                                                                      \
        if (!(cond)) return efail((e), FP_E_STATE, (msg));                                    \
    } while (0)

template <typename T>
hipError_t dmalloc(T **p, uint64_t count)
{
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

void drop_graph_slot(fp_ens *e, int s)
{
    if (e->exec[s]) (void)hipGraphExecDestroy(e->exec[s]);
    if (e->graph[s]) (void)hipGraphDestroy(e->graph[s]);
    e->exec[s] = nullptr;
    e->graph[s] = nullptr;
    if (s == 1) e->multi_steps = 0;
}

void drop_graph(fp_ens *e)
{
    drop_graph_slot(e, 0);
    drop_graph_slot(e, 1);
}

// Capture n_steps consecutive sync steps (clock + step launch each) into
// graph slot s.  The clock lives on the device, so the replays need no host
// arguments.
int capture_steps(fp_ens *e, int s, uint32_t n_steps)
{
    drop_graph_slot(e, s);
    E_HIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    hipError_t h = hipSuccess;
    for (uint32_t k = 0; k < n_steps && h == hipSuccess; ++k)
        h = fp::launch_ens_step(e->d_members, e->d_blocks, e->d_clock, e->nblocks, e->block, e->ppt, e->sub4,
                                e->soa.capacity <= fp::kSoa32Slots, e->stream);
    hipGraph_t g = nullptr;
    const hipError_t h2 = hipStreamEndCapture(e->stream, &g);
    if (h != hipSuccess || h2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        E_HIP(e, h);
        E_HIP(e, h2);
    }
    e->graph[s] = g;
    E_HIP(e, hipGraphInstantiate(&e->exec[s], e->graph[s], nullptr, nullptr, 0));
    e->graph_sig[s] = e->sig();
    if (s == 1) e->multi_steps = n_steps;
    return FP_OK;
}

void destroy(fp_ens *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    auto fr = [](void *p) { if (p) (void)hipFree(p); };
    fr(e->soa.base); fr(e->met); fr(e->met_stage); fr(e->level_heights); fr(e->pbl); fr(e->precip); fr(e->pbl_stage);
    fr(e->grid_counts); fr(e->grid_mass); fr(e->aos_stage); fr(e->d_members); fr(e->d_blocks);
    fr(e->d_clock);
    fp::sort_work_free(e->sortw);
    if (e->ev_a) (void)hipEventDestroy(e->ev_a);
    if (e->ev_b) (void)hipEventDestroy(e->ev_b);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int met_member_common(fp_ens *e, uint32_t m, const float *const f[6], hipMemcpyKind kind, const float *lh)
{
    E_CHECK(e, m < e->n_members, "member index out of range");
    for (uint32_t k = 0; k < e->nz; ++k) E_CHECK(e, std::isfinite(lh[k]), "level heights must be finite");
    const bool metres = lh[e->nz - 1] > 0.0f;
    if (metres)
        for (uint32_t k = 1; k < e->nz; ++k) E_CHECK(e, lh[k] >= lh[k - 1], "level heights must be non-decreasing");
    bool any = false;
    for (char r : e->met_ready) any = any || r;
    if (any) E_CHECK(e, std::equal(lh, lh + e->nz, e->lh_host.begin()), "members must share the level heights");
    e->has_heights = fp::level_table(lh, e->nz, e->lh_host);
    e->zmax = metres ? lh[e->nz - 1] : (float)(e->nz - 1);
    E_HIP(e, hipMemcpyAsync(e->level_heights, e->lh_host.data(), e->lh_host.size() * 4, hipMemcpyHostToDevice,
                            e->stream));
    const uint64_t np = e->npts();
    const float *src[6];
    if (kind == hipMemcpyHostToDevice) {
        // stage the six host fields, then pack on the device
        if (!e->met_stage) E_HIP(e, dmalloc(&e->met_stage, np * 6));
        for (int k = 0; k < 6; ++k) {
            E_HIP(e, hipMemcpyAsync(e->met_stage + k * np, f[k], np * 4, hipMemcpyHostToDevice, e->stream));
            src[k] = e->met_stage + k * np;
        }
    } else {
        for (int k = 0; k < 6; ++k) src[k] = f[k];
    }
    E_HIP(e, fp::launch_pack_met(src[0], src[1], src[2], src[3], src[4], src[5], e->met + (uint64_t)m * np * 2,
                                 (uint32_t)np, e->stream));
    E_HIP(e, hipStreamSynchronize(e->stream));
    e->met_ready[m] = 1;
    return FP_OK;
}

// Launch tables for the current block size and particles per thread from the
// configured members (the clock is left as it is).
int rebuild_tables(fp_ens *e)
{
    std::vector<fp::EnsBlock> blocks;
    const uint64_t per_block = (uint64_t)e->block * e->ppt;
    for (uint32_t m = 0; m < e->n_members; ++m) {
        e->members[m].a.ppt = e->ppt;
        for (uint64_t f = 0; f < e->members[m].a.n; f += per_block) blocks.push_back(fp::EnsBlock{m, (uint32_t)f});
    }
    e->nblocks = (uint32_t)blocks.size();
    if (blocks.size() > e->d_blocks_cap) {
        if (e->d_blocks) {
            E_HIP(e, hipStreamSynchronize(e->stream));
            (void)hipFree(e->d_blocks);
        }
        e->d_blocks = nullptr;
        e->d_blocks_cap = 0;
        E_HIP(e, dmalloc(&e->d_blocks, blocks.size()));
        e->d_blocks_cap = blocks.size();
    }
    E_HIP(e, hipMemcpyAsync(e->d_members, e->members.data(), e->members.size() * sizeof(fp::EnsMember),
                            hipMemcpyHostToDevice, e->stream));
    if (!blocks.empty())
        E_HIP(e, hipMemcpyAsync(e->d_blocks, blocks.data(), blocks.size() * sizeof(fp::EnsBlock),
                                hipMemcpyHostToDevice, e->stream));
    E_HIP(e, hipStreamSynchronize(e->stream));
    return FP_OK;
}

}  // namespace

extern "C" {

const char *fp_ens_last_error(const fp_ens *e) { return e ? e->err.c_str() : g_ens_error.c_str(); }

int fp_ens_create(int device_ordinal, const fp_ens_desc *desc, fp_ens **out)
{
    if (!desc || !out || !desc->capacities) return efail(nullptr, FP_E_INVALID, "null argument");
    *out = nullptr;
    const fp_ens_desc &d = *desc;
    if (d.n_members == 0) return efail(nullptr, FP_E_INVALID, "n_members must be >= 1");
    if (d.met_nx == 0 || d.met_ny == 0 || d.met_nz == 0 || d.met_nz > fp::kMaxLevels)
        return efail(nullptr, FP_E_INVALID, "met grid dimensions must be in [1, 256] levels");
    if (d.n_species < 1 || d.n_species > FP_MAX_SPECIES) return efail(nullptr, FP_E_INVALID, "n_species must be 1..4");
    if (d.out_nz > FP_MAX_OUTPUT_LEVELS) return efail(nullptr, FP_E_INVALID, "out_nz exceeds 16");
    uint64_t total = 0, mx = 0;
    for (uint32_t m = 0; m < d.n_members; ++m) {
        total += d.capacities[m];
        mx = std::max<uint64_t>(mx, d.capacities[m]);
    }
    if (total == 0 || total > fp::kMaxSlots) return efail(nullptr, FP_E_INVALID, "total capacity must be in [1, 2^30]");
    const uint64_t npts = (uint64_t)d.met_nx * d.met_ny * d.met_nz;
    if (npts > fp::kMaxMetPoints) return efail(nullptr, FP_E_INVALID, "met grid too large (at most 2^27 points)");
    if ((uint64_t)d.out_nx * d.out_ny * d.out_nz >= 0x80000000ull)
        return efail(nullptr, FP_E_INVALID, "output grid too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return efail(nullptr, FP_E_DEVICE, "no HIP device");
    if (device_ordinal < 0 || device_ordinal >= ndev) return efail(nullptr, FP_E_INVALID, "device ordinal out of range");
    if (hipSetDevice(device_ordinal) != hipSuccess) return efail(nullptr, FP_E_DEVICE, "hipSetDevice failed");

    fp_ens *e = new (std::nothrow) fp_ens();
    if (!e) return efail(nullptr, FP_E_NOMEM, "host allocation failed");
    e->device = device_ordinal;
    e->n_members = d.n_members;
    e->nx = d.met_nx; e->ny = d.met_ny; e->nz = d.met_nz;
    e->n_species = d.n_species;
    e->ox = d.out_nx; e->oy = d.out_ny; e->oz = d.out_nz;
    e->cap.assign(d.capacities, d.capacities + d.n_members);
    e->off.resize(d.n_members);
    uint64_t acc = 0;
    for (uint32_t m = 0; m < d.n_members; ++m) { e->off[m] = acc; acc += e->cap[m]; }
    e->total = total;
    e->max_cap = mx;
    e->met_ready.assign(d.n_members, 0);
    auto bail = [&](hipError_t h, const char *what) {
        g_ens_error = std::string(what) + ": " + hipGetErrorString(h);
        destroy(e);
        return h == hipErrorOutOfMemory ? FP_E_NOMEM : FP_E_DEVICE;
    };
#define C_HIP(expr)                                  \
    do {                                             \
        hipError_t h_ = (expr);                      \
        if (h_ != hipSuccess) return bail(h_, #expr); \
    } while (0)
    C_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    e->soa.capacity = total;
    C_HIP(dmalloc(&e->soa.base, total * fp::F_COUNT));
    C_HIP(hipMemsetAsync(e->soa.base, 0, total * fp::F_COUNT * 4, e->stream));
    C_HIP(dmalloc(&e->met, npts * 2 * d.n_members));
    C_HIP(dmalloc(&e->level_heights, (uint64_t)fp::level_table_words(d.met_nz)));
    const uint64_t np = (uint64_t)d.met_nx * d.met_ny;
    C_HIP(dmalloc(&e->pbl, np * 2));
    C_HIP(dmalloc(&e->precip, np));
    C_HIP(dmalloc(&e->pbl_stage, np * 4));
    if (e->ncell()) {
        C_HIP(dmalloc(&e->grid_counts, e->ncell() * d.n_members));
        C_HIP(dmalloc(&e->grid_mass, e->ncell() * d.n_members));
        C_HIP(hipMemsetAsync(e->grid_counts, 0, e->ncell() * d.n_members * 4, e->stream));
        C_HIP(hipMemsetAsync(e->grid_mass, 0, e->ncell() * d.n_members * 4, e->stream));
    }
    C_HIP(dmalloc(&e->aos_stage, std::max<uint64_t>(mx, 1)));
    C_HIP(dmalloc(&e->d_members, (uint64_t)d.n_members));
    C_HIP(dmalloc(&e->d_clock, 1));
    C_HIP(hipEventCreate(&e->ev_a));
    C_HIP(hipEventCreate(&e->ev_b));
    C_HIP(hipStreamSynchronize(e->stream));
#undef C_HIP
    e->gp.refine_x = e->gp.refine_y = 1;
    e->gp.mass_scale = 1.0e6f;
    *out = e;
    return FP_OK;
}

int fp_ens_destroy(fp_ens *e)
{
    destroy(e);
    return FP_OK;
}

int fp_ens_sync(fp_ens *e)
{
    E_CHECK(e, e, "null ensemble");
    E_HIP(e, hipStreamSynchronize(e->stream));
    return FP_OK;
}

int fp_ens_upload_member(fp_ens *e, uint32_t m, const fp_particle *p, uint64_t n)
{
    E_CHECK(e, e && (p || n == 0), "null argument");
    E_CHECK(e, m < e->n_members && n <= e->cap[m], "member index or count out of range");
    for (uint64_t r = 0; r < n; ++r)
        for (uint32_t s = e->n_species; s < FP_MAX_SPECIES; ++s)
            E_CHECK(e, p[r].mass[s] == 0.0f, "mass slot beyond n_species is non-zero");
    if (n == 0) return FP_OK;
    fp::Soa s = e->soa;
    s.base += e->off[m];
    E_HIP(e, hipMemcpyAsync(e->aos_stage, p, n * sizeof(fp_particle), hipMemcpyHostToDevice, e->stream));
    E_HIP(e, fp::launch_aos_to_soa(e->aos_stage, nullptr, n, s, e->stream));
    E_HIP(e, hipStreamSynchronize(e->stream));
    return FP_OK;
}

int fp_ens_download_member(fp_ens *e, uint32_t m, fp_particle *out, uint64_t n)
{
    E_CHECK(e, e && (out || n == 0), "null argument");
    E_CHECK(e, m < e->n_members && n <= e->cap[m], "member index or count out of range");
    if (n == 0) return FP_OK;
    fp::Soa s = e->soa;
    s.base += e->off[m];
    E_HIP(e, fp::launch_soa_to_aos(s, e->aos_stage, n, e->stream));
    E_HIP(e, hipMemcpyAsync(out, e->aos_stage, n * sizeof(fp_particle), hipMemcpyDeviceToHost, e->stream));
    E_HIP(e, hipStreamSynchronize(e->stream));
    return FP_OK;
}

int fp_ens_met_upload_member(fp_ens *e, uint32_t m, const float *u0, const float *v0, const float *w0,
                             const float *u1, const float *v1, const float *w1, const float *lh)
{
    E_CHECK(e, e && u0 && v0 && w0 && u1 && v1 && w1 && lh, "null argument");
    const float *f[6] = {u0, v0, w0, u1, v1, w1};
    return met_member_common(e, m, f, hipMemcpyHostToDevice, lh);
}

int fp_ens_met_upload_member_dev(fp_ens *e, uint32_t m, const float *u0, const float *v0, const float *w0,
                                 const float *u1, const float *v1, const float *w1, const float *lh)
{
    E_CHECK(e, e && u0 && v0 && w0 && u1 && v1 && w1 && lh, "null argument");
    const float *f[6] = {u0, v0, w0, u1, v1, w1};
    return met_member_common(e, m, f, hipMemcpyDeviceToDevice, lh);
}

int fp_ens_pbl_upload(fp_ens *e, const float *ustar, const float *wstar, const float *hmix,
                      const float *oli, const float *precip)
{
    E_CHECK(e, e && ustar && wstar && hmix && oli, "null argument");
    const uint64_t np = (uint64_t)e->nx * e->ny;
    const float *f[4] = {ustar, wstar, hmix, oli};
    for (int k = 0; k < 4; ++k)
        E_HIP(e, hipMemcpyAsync(e->pbl_stage + k * np, f[k], np * 4, hipMemcpyHostToDevice, e->stream));
    E_HIP(e, fp::launch_pack_pbl(e->pbl_stage, e->pbl, (uint32_t)np, e->stream));
    if (precip) {
        E_HIP(e, hipMemcpyAsync(e->precip, precip, np * 4, hipMemcpyHostToDevice, e->stream));
        e->precip_ready = true;
    }
    E_HIP(e, hipStreamSynchronize(e->stream));
    e->pbl_ready = true;
    return FP_OK;
}

int fp_ens_grid_configure(fp_ens *e, const fp_grid_params *p)
{
    E_CHECK(e, e && p, "null argument");
    E_CHECK(e, e->grid_counts, "ensemble has no output grid");
    E_CHECK(e, p->refine_x >= 1 && p->refine_y >= 1, "refine factors must be >= 1");
    E_CHECK(e, p->species < e->n_species, "species index beyond n_species");
    E_CHECK(e, std::isfinite(p->mass_scale), "mass_scale must be finite");
    e->gp = *p;
    e->grid_configured = true;
    e->configured = false;  // launch tables hold the grid arguments
    return FP_OK;
}

int fp_ens_grid_clear(fp_ens *e)
{
    E_CHECK(e, e && e->grid_counts, "ensemble has no output grid");
    const uint64_t n = e->ncell() * e->n_members * 4;
    E_HIP(e, hipMemsetAsync(e->grid_counts, 0, n, e->stream));
    E_HIP(e, hipMemsetAsync(e->grid_mass, 0, n, e->stream));
    return FP_OK;
}

int fp_ens_grid_download_member(fp_ens *e, uint32_t m, uint32_t *counts, uint32_t *mass)
{
    E_CHECK(e, e && e->grid_counts, "ensemble has no output grid");
    E_CHECK(e, m < e->n_members, "member index out of range");
    const uint64_t nc = e->ncell();
    if (counts)
        E_HIP(e, hipMemcpyAsync(counts, e->grid_counts + m * nc, nc * 4, hipMemcpyDeviceToHost, e->stream));
    if (mass)
        E_HIP(e, hipMemcpyAsync(mass, e->grid_mass + m * nc, nc * 4, hipMemcpyDeviceToHost, e->stream));
    E_HIP(e, hipStreamSynchronize(e->stream));
    return FP_OK;
}

int fp_ens_set_launch(fp_ens *e, uint32_t block, uint32_t ppt)
{
    E_CHECK(e, e, "null ensemble");
    E_CHECK(e, block == 64 || block == 128 || block == 256 || block == 512 || block == 1024,
            "block_size must be 64, 128, 256, 512 or 1024");
    E_CHECK(e, ppt >= 1 && ppt <= 16, "particles_per_thread must be 1..16");
    e->block = block;
    e->ppt = ppt;
    // a configured ensemble keeps its configuration and clock
    if (e->configured) return rebuild_tables(e);
    return FP_OK;
}

int fp_ens_configure(fp_ens *e, const fp_step_params *p, const uint32_t *keys, const uint32_t *counters,
                     const uint64_t *dispatch, double t0_in_bracket, double bracket_s)
{
    E_CHECK(e, e && p, "null argument");
    E_STATE(e, e->pbl_ready, "fp_ens_configure before fp_ens_pbl_upload");
    for (uint32_t m = 0; m < e->n_members; ++m) E_STATE(e, e->met_ready[m], "member met bracket missing");
    E_CHECK(e, std::isfinite(p->dt_seconds) && p->dt_seconds > 0.0f, "invalid dt_seconds");
    E_CHECK(e, p->n_substeps <= 4, "n_substeps must be 0..4");
    E_CHECK(e, p->direction == 1 || p->direction == -1, "direction must be +1 or -1");
    E_CHECK(e, p->dry_mode <= FP_DRY_UNIFORM, "ensembles support dry_mode off or uniform");
    E_CHECK(e, p->wet_mode != FP_WET_PER_SLOT && p->wet_mode <= FP_WET_PRECIP_FIELD, "unsupported wet_mode");
    E_STATE(e, p->wet_mode != FP_WET_PRECIP_FIELD || e->precip_ready, "precip field missing");
    E_CHECK(e, p->loss_mode <= FP_LOSS_MIN_MASS, "unknown loss_mode");
    E_STATE(e, !p->grid_fused || e->grid_configured, "fused gridding before fp_ens_grid_configure");
    E_CHECK(e, bracket_s > 0.0 && std::isfinite(t0_in_bracket), "invalid bracket time");
    e->members.assign(e->n_members, fp::EnsMember{});
    const uint64_t np = e->npts();
    for (uint32_t m = 0; m < e->n_members; ++m) {
        const uint64_t n = dispatch ? dispatch[m] : e->cap[m];
        E_CHECK(e, n <= e->cap[m], "dispatch exceeds member capacity");
        fp::StepArgs &a = e->members[m].a;
        a.s = e->soa;
        a.s.base += e->off[m];
        a.n = (uint32_t)n;
        a.n_species = e->n_species;
        a.met = e->met + (uint64_t)m * np * 2;
        a.nx = e->nx; a.ny = e->ny; a.nz = e->nz;
        a.level_heights = e->level_heights;
        a.has_heights = e->has_heights;
        a.zmax = e->zmax;
        a.pbl = e->pbl;
        a.precip = e->precip;
        fp::fill_step_scalars(a, *p);
        a.alpha = 0.0f;  // the ensemble clock supplies the temporal weight
        if (keys) {
            a.key0 = keys[2 * m];
            a.key1 = keys[2 * m + 1];
        }
        if (e->grid_counts)
            a.grid = fp::make_grid_args(e->grid_counts + (uint64_t)m * e->ncell(),
                                        e->grid_mass + (uint64_t)m * e->ncell(), e->ox, e->oy, e->oz, e->gp,
                                        e->nx, e->ny);
        a.ppt = e->ppt;
        for (int k = 0; k < 4; ++k) e->members[m].base_ctr[k] = counters ? counters[4 * m + k] : p->counter[k];
        e->members[m].ctr_per_step = fp::counter_per_step(n, p->n_substeps);
    }
    if (int rc = rebuild_tables(e)) return rc;
    fp::EnsClock clk{};
    clk.next_step = 0;
    clk.t0_in_bracket = t0_in_bracket;
    clk.dt_seconds = p->direction < 0 ? -(double)p->dt_seconds : (double)p->dt_seconds;
    clk.bracket_seconds = bracket_s;
    E_HIP(e, hipMemcpyAsync(e->d_clock, &clk, sizeof clk, hipMemcpyHostToDevice, e->stream));
    E_HIP(e, hipStreamSynchronize(e->stream));
    e->sub4 = p->n_substeps == 4;
    e->steps_done = 0;
    e->configured = true;
    return FP_OK;
}

int fp_ens_step(fp_ens *e, uint32_t n_steps, int use_graph, uint32_t *counters_out)
{
    E_CHECK(e, e, "null ensemble");
    E_STATE(e, e->configured, "fp_ens_step before fp_ens_configure");
    if (n_steps == 0) return FP_OK;
    // graph path: the call's steps as replays of one captured sequence of
    // min(n_steps, kMaxGraphSteps) steps, kept for later calls of that
    // length, and the remainder from the one-step graph
    constexpr uint32_t kMaxGraphSteps = 36;
    const uint32_t group = std::min(n_steps, kMaxGraphSteps);
    const uint32_t n_groups = group > 1 ? n_steps / group : 0;
    const uint32_t n_single = n_steps - n_groups * group;
    if (use_graph) {
        const auto now = e->sig();
        if (n_groups && !(e->exec[1] && e->multi_steps == group && e->graph_sig[1] == now))
            if (int rc = capture_steps(e, 1, group)) return rc;
        if (n_single && !(e->exec[0] && e->graph_sig[0] == now))
            if (int rc = capture_steps(e, 0, 1)) return rc;
    }
    E_HIP(e, hipEventRecord(e->ev_a, e->stream));
    if (use_graph) {
        for (uint32_t k = 0; k < n_groups; ++k) E_HIP(e, hipGraphLaunch(e->exec[1], e->stream));
        for (uint32_t k = 0; k < n_single; ++k) E_HIP(e, hipGraphLaunch(e->exec[0], e->stream));
    } else {
        for (uint32_t k = 0; k < n_steps; ++k)
            E_HIP(e, fp::launch_ens_step(e->d_members, e->d_blocks, e->d_clock, e->nblocks, e->block, e->ppt,
                                         e->sub4, e->soa.capacity <= fp::kSoa32Slots, e->stream));
    }
    E_HIP(e, hipEventRecord(e->ev_b, e->stream));
    E_HIP(e, hipEventSynchronize(e->ev_b));
    float ms = 0.0f;
    E_HIP(e, hipEventElapsedTime(&ms, e->ev_a, e->ev_b));
    e->last_ms = ms / n_steps;
    e->steps_done += n_steps;
    if (counters_out)
        for (uint32_t m = 0; m < e->n_members; ++m)
            fp::counter_advance(e->members[m].base_ctr, e->steps_done * e->members[m].ctr_per_step,
                                counters_out + 4 * m);
    return FP_OK;
}

int fp_ens_last_step_ms(fp_ens *e, double *ms)
{
    E_CHECK(e, e && ms, "null argument");
    *ms = e->last_ms;
    return FP_OK;
}

// Morton re-sort of every member's dispatch range (the member's capacity
// before the first fp_ens_configure): the order fp_sort_morton gives for the
// member alone.  Members are sorted in place through the AoS stage.
int fp_ens_sort_morton(fp_ens *e, const fp_sort_params *p)
{
    E_CHECK(e, e && p, "null argument");
    E_CHECK(e, p->bits_per_axis >= 1 && p->bits_per_axis <= 21, "bits_per_axis must be 1..21");
    if (!e->sortw.cap) {
        const hipError_t h = fp::sort_work_alloc(e->sortw, e->max_cap);
        if (h != hipSuccess) {
            fp::sort_work_free(e->sortw);
            E_HIP(e, h);
        }
    }
    for (uint32_t m = 0; m < e->n_members; ++m) {
        const uint64_t n = e->members.empty() ? e->cap[m] : e->members[m].a.n;
        if (n == 0) continue;
        fp::Soa s = e->soa;
        s.base += e->off[m];
        E_HIP(e, fp::sort_morton_range(e->sortw, *p, s, s, e->aos_stage, (uint32_t)n, e->stream));
    }
    E_HIP(e, hipStreamSynchronize(e->stream));
    return FP_OK;
}

// Launch shape and re-sort interval of the ensemble's graph path: block size x
// particles per thread timed on the current state (warm-up, then the median of
// measured steps), then the re-sort intervals with the winning shape, timed
// with the sorts amortised; ties keep the earlier candidate (the reference's
// workgroup autotuner, gpu/workgroup.rs:342-445).  The particle state, the
// grids and the step clock are restored afterwards, so the run continues
// exactly as without the call.  params->cache_path is not used.
int fp_ens_autotune(fp_ens *e, const fp_autotune_params *params, fp_autotune_result *out)
{
    E_CHECK(e, e && params && out, "null argument");
    E_STATE(e, e->configured, "fp_ens_autotune before fp_ens_configure");
    E_CHECK(e, params->measured_steps >= 1 && params->n_sort_intervals <= 8,
            "measured_steps must be >= 1 and n_sort_intervals <= 8");
    std::memset(out, 0, sizeof *out);
    const uint64_t words = e->total * fp::F_COUNT;
    const uint64_t gwords = e->grid_counts ? e->ncell() * e->n_members : 0;
    uint32_t *snap = nullptr;
    E_HIP(e, dmalloc(&snap, words + 2 * gwords));
    fp::EnsClock clk{};
    auto copy_all = [&](bool save) -> hipError_t {
        hipError_t h = hipMemcpyAsync(save ? snap : e->soa.base, save ? e->soa.base : snap, words * 4,
                                      hipMemcpyDeviceToDevice, e->stream);
        if (h == hipSuccess && gwords)
            h = hipMemcpyAsync(save ? snap + words : e->grid_counts, save ? e->grid_counts : snap + words,
                               gwords * 4, hipMemcpyDeviceToDevice, e->stream);
        if (h == hipSuccess && gwords)
            h = hipMemcpyAsync(save ? snap + words + gwords : e->grid_mass,
                               save ? e->grid_mass : snap + words + gwords, gwords * 4, hipMemcpyDeviceToDevice,
                               e->stream);
        if (h == hipSuccess)
            h = save ? hipMemcpyAsync(&clk, e->d_clock, sizeof clk, hipMemcpyDeviceToHost, e->stream)
                     : hipMemcpyAsync(e->d_clock, &clk, sizeof clk, hipMemcpyHostToDevice, e->stream);
        if (h == hipSuccess) h = hipStreamSynchronize(e->stream);
        return h;
    };
    const uint32_t block0 = e->block, ppt0 = e->ppt;
    const uint64_t steps0 = e->steps_done;
    int rc = FP_OK;
    hipEvent_t ta = nullptr, tb = nullptr;  // fp_ens_step records ev_a/ev_b itself
    hipError_t h = hipEventCreate(&ta);
    if (h == hipSuccess) h = hipEventCreate(&tb);
    if (h == hipSuccess) h = copy_all(true);
    auto timed = [&](uint32_t steps, uint32_t sort_every, std::vector<double> *per_step, double *mean) -> int {
        fp_sort_params sp{};
        sp.bits_per_axis = params->sort_bits ? params->sort_bits : 10;
        double total = 0.0;
        for (uint32_t k = 0; k < steps; ++k) {
            E_HIP(e, hipEventRecord(ta, e->stream));
            if (sort_every && k % sort_every == 0)
                if (int r = fp_ens_sort_morton(e, &sp)) return r;
            if (int r = fp_ens_step(e, 1, 1, nullptr)) return r;
            E_HIP(e, hipEventRecord(tb, e->stream));
            E_HIP(e, hipEventSynchronize(tb));
            float ms = 0.0f;
            E_HIP(e, hipEventElapsedTime(&ms, ta, tb));
            if (per_step) per_step->push_back(ms);
            total += ms;
        }
        if (mean) *mean = steps ? total / steps : 0.0;
        return FP_OK;
    };
    static const uint32_t blocks[] = {128, 256, 512};
    static const uint32_t ppts[] = {1, 2, 4};
    double best = 1e300;
    uint32_t best_b = block0, best_p = ppt0, best_s = 0;
    if (h != hipSuccess) goto done;
    for (uint32_t b : blocks)
        for (uint32_t pp : ppts) {
            if ((rc = fp_ens_set_launch(e, b, pp))) goto done;
            if ((h = copy_all(false)) != hipSuccess) goto done;
            if ((rc = timed(params->warmup_steps, 0, nullptr, nullptr))) goto done;
            std::vector<double> t;
            if ((rc = timed(params->measured_steps, 0, &t, nullptr))) goto done;
            std::sort(t.begin(), t.end());
            if (t[t.size() / 2] < best) {
                best = t[t.size() / 2];
                best_b = b;
                best_p = pp;
            }
        }
    out->step_ms = best;
    out->step_ms_with_sort = best;
    if ((rc = fp_ens_set_launch(e, best_b, best_p))) goto done;
    for (uint32_t c = 0; c < params->n_sort_intervals; ++c) {
        const uint32_t every = params->sort_intervals[c];
        const uint32_t steps = std::max<uint32_t>(params->sort_trial_steps, 1u);
        if ((h = copy_all(false)) != hipSuccess) goto done;
        double mean = 0.0;
        if ((rc = timed(steps, every, nullptr, &mean))) goto done;
        if (c == 0 || mean < out->step_ms_with_sort) {
            out->step_ms_with_sort = mean;
            best_s = every;
        }
    }
done:
    {
        // the state, grids and clock as before the call, with the chosen shape
        // (or the original one after an error)
        const int rc2 = fp_ens_set_launch(e, rc == FP_OK && h == hipSuccess ? best_b : block0,
                                          rc == FP_OK && h == hipSuccess ? best_p : ppt0);
        const hipError_t h2 = copy_all(false);
        e->steps_done = steps0;
        (void)hipFree(snap);
        if (ta) (void)hipEventDestroy(ta);
        if (tb) (void)hipEventDestroy(tb);
        E_HIP(e, h);
        if (rc) return rc;
        if (rc2) return rc2;
        E_HIP(e, h2);
    }
    out->block_size = best_b;
    out->particles_per_thread = best_p;
    out->sort_interval = best_s;
    return FP_OK;
}

}  // extern "C"
