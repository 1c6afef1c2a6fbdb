// fp_release.hip - source-term injection on the device (fp_release_inject):
// ReleaseManager::inject_for_time (release/mod.rs:90-152) with the particles
// generated where they are stored instead of uploaded record by record.
//
// The host decides, per release, how many particles the time window emits by
// now (target_emitted_count, :181-197) and checks every sample against the
// grid domain, so that errors leave the store exactly as the reference leaves
// it; the device then places each particle (sample_stratified, :264-283,
// geo_to_grid and to_relative, coords/mod.rs:109-114,169-178) in binary64,
// with the reference's operation order, and writes its 24 SoA words.
#include <cmath>
#include <string>
#include <vector>

#include "fp_internal.hpp"

namespace fp {
namespace {

constexpr uint32_t kReleaseBlock = 256;

struct Strata { uint64_t nx, nxy, nz; };

// the strata of sample_stratified for a release of `count` particles
Strata strata_of(uint64_t count)
{
    double c = std::ceil(std::cbrt((double)count));
    if (c < 1.0) c = 1.0;
    Strata s;
    s.nx = (uint64_t)c;
    s.nxy = s.nx * s.nx;
    if (s.nxy < 1) s.nxy = 1;
    s.nz = (count + s.nxy - 1) / s.nxy;
    if (s.nz < 1) s.nz = 1;
    return s;
}

struct Sample { double lon, lat, z; };

__host__ __device__ inline Sample sample_stratified(const fp_release &r, const Strata &st, uint64_t idx)
{
    const uint64_t ix = idx % st.nx, iy = (idx / st.nx) % st.nx, iz = (idx / st.nxy) % st.nz;
    const double fx = ((double)ix + 0.5) / (double)st.nx;
    const double fy = ((double)iy + 0.5) / (double)st.nx;
    const double fz = ((double)iz + 0.5) / (double)st.nz;
    Sample s;
    s.lon = r.lon_min + fx * (r.lon_max - r.lon_min);
    s.lat = r.lat_min + fy * (r.lat_max - r.lat_min);
    s.z = r.z_min + fz * (r.z_max - r.z_min);
    return s;
}

// One release's share of an injection.
struct Job {
    fp_release r;
    Strata st;
    uint64_t first_index;  // sample index of its first particle
    uint64_t out_first;    // its first particle's position in the injection order
    float mass;
    uint32_t pad;
};

__global__ void __launch_bounds__(kReleaseBlock) release_kernel(const Job *jobs, uint32_t n_jobs, uint64_t total,
                                                                fp_grid_domain g, int32_t time, const uint32_t *slots,
                                                                uint64_t first_slot, Soa s)
{
    const uint64_t t = (uint64_t)blockIdx.x * kReleaseBlock + threadIdx.x;
    if (t >= total) return;
    uint32_t lo = 0, hi = n_jobs;  // last job with out_first <= t
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].out_first <= t) lo = mid; else hi = mid;
    }
    const Job &j = jobs[lo];
    const Sample p = sample_stratified(j.r, j.st, j.first_index + (t - j.out_first));
    const double gx = (p.lon - g.xlon0) / g.dx;
    const double gy = (p.lat - g.ylat0) / g.dy;
    const double fx = floor(gx), fy = floor(gy);
    const uint64_t i = slots ? (uint64_t)slots[t] : first_slot + t;
    for (uint32_t f = 0; f < F_COUNT; ++f) s.u(f)[i] = 0u;
    s.f(F_POS_X)[i] = (float)(gx - fx);
    s.f(F_POS_Y)[i] = (float)(gy - fy);
    s.f(F_POS_Z)[i] = (float)p.z;
    s.i(F_CELL_X)[i] = (int32_t)fx;
    s.i(F_CELL_Y)[i] = (int32_t)fy;
    s.u(F_FLAGS)[i] = 1u;  // FLAG_ACTIVE
    s.f(F_MASS0)[i] = j.mass;
    s.i(F_TIME)[i] = time;
    s.i(F_TIME_MEM)[i] = time;
    s.i(F_RELEASE_POINT)[i] = j.r.release_point;
    s.i(F_CLASS_ID)[i] = j.r.class_id;
}

// target_emitted_count (release/mod.rs:181-197)
uint64_t target_count(const fp_release &r, uint64_t released, int64_t now)
{
    if (now < r.start_s || now > r.end_s) return released;
    if (r.start_s == r.end_s || now == r.end_s) return r.particle_count;
    const uint64_t elapsed = (uint64_t)(now - r.start_s), duration = (uint64_t)(r.end_s - r.start_s);
    return (uint64_t)(((unsigned __int128)r.particle_count * elapsed) / duration);
}

}  // namespace

}  // namespace fp

namespace fp {

// Plans an injection on the host and writes its particles on the device.
// *status: 0, 1 (a sample outside the grid domain: the particles before it
// are written and that release's count is kept) or 2 (out of free slots).
hipError_t release_plan_and_launch(const fp_grid_domain &g, const fp_release *rel, uint32_t n_rel,
                                   uint64_t *released, int64_t now, int32_t time, const uint32_t *slots_dev,
                                   uint64_t first_slot, uint64_t max_particles, const Soa &s, void **jobs_dev,
                                   uint64_t *jobs_cap_bytes, hipStream_t st, uint64_t *injected, int *status)
{
    std::vector<Job> jobs;
    uint64_t total = 0;
    *status = 0;
    for (uint32_t k = 0; k < n_rel && *status == 0; ++k) {
        const fp_release &r = rel[k];
        const uint64_t target = target_count(r, released[k], now);
        if (target <= released[k]) continue;
        Job j{};
        j.r = r;
        j.st = strata_of(r.particle_count);
        j.first_index = released[k];
        j.out_first = total;
        j.mass = (float)(r.mass_kg / (double)r.particle_count);
        uint64_t idx = released[k];
        // Every sample's grid x depends only on its x stratum and grows (or
        // falls) monotonically with it in floating point (the stratum centre,
        // the scaled extent, the offset and the division by dx are each
        // monotone), and the same holds for y; so when the four extreme
        // strata corners map inside the domain, every sample does and the
        // per-sample check below can only stop at the slot limit.
        bool inside = true;
        for (uint64_t cx : {(uint64_t)0, j.st.nx - 1u})
            for (uint64_t cy : {(uint64_t)0, j.st.nx - 1u}) {
                const Sample p = sample_stratified(r, j.st, cx + cy * j.st.nx);
                const double gx = (p.lon - g.xlon0) / g.dx, gy = (p.lat - g.ylat0) / g.dy;
                inside = inside && gx >= 0.0 && gx < (double)(g.nx - 1u) && gy >= 0.0 && gy < (double)(g.ny - 1u);
            }
        if (inside) {
            const uint64_t room = max_particles > total ? max_particles - total : 0u;
            if (target - idx > room) {
                idx += room;
                *status = 2;
            } else {
                idx = target;
            }
        }
        for (; *status == 0 && idx < target; ++idx) {
            const Sample p = sample_stratified(r, j.st, idx);
            const double gx = (p.lon - g.xlon0) / g.dx, gy = (p.lat - g.ylat0) / g.dy;
            if (!(gx >= 0.0 && gx < (double)(g.nx - 1u) && gy >= 0.0 && gy < (double)(g.ny - 1u))) {
                *status = 1;
                break;
            }
            if (total + (idx - released[k]) >= max_particles) {
                *status = 2;
                break;
            }
        }
        const uint64_t emitted = idx - released[k];
        if (emitted) jobs.push_back(j);
        total += emitted;
        if (*status == 0) released[k] = target;
    }
    *injected = total;
    if (total == 0) return hipSuccess;
    const uint64_t bytes = jobs.size() * sizeof(Job);
    if (bytes > *jobs_cap_bytes) {
        if (*jobs_dev) (void)hipFree(*jobs_dev);
        *jobs_dev = nullptr;
        *jobs_cap_bytes = 0;
        if (hipError_t h = hipMalloc(jobs_dev, bytes)) return h;
        *jobs_cap_bytes = bytes;
    }
    if (hipError_t h = hipMemcpyAsync(*jobs_dev, jobs.data(), bytes, hipMemcpyHostToDevice, st)) return h;
    const uint32_t grid = (uint32_t)((total + kReleaseBlock - 1) / kReleaseBlock);
    release_kernel<<<grid, kReleaseBlock, 0, st>>>(reinterpret_cast<const Job *>(*jobs_dev), (uint32_t)jobs.size(),
                                                   total, g, time, slots_dev, first_slot, s);
    if (hipError_t h = hipGetLastError()) return h;
    return hipStreamSynchronize(st);
}

}  // namespace fp
