// fp_output.hip - concentration output: interval sums on the device, the
// reference's kg conversion, time means and cell-volume normalisation, and
// the NetCDF snapshot file.
//
// The reference downloads the fixed-point grid at each output time and
// converts it on the host (accumulate_concentration_grid_gpu,
// gpu/gridding.rs:194-219: mass_kg = (value as f32) / mass_scale), and the
// science notes define the concentration as C = mass / (mass_scale * V_cell),
// V_cell = dx_m * dy_m * dz_m (docs/science/concentration-gridding.md:35-41).
// Here every sampled step adds the grid to 64-bit interval sums in HBM (no
// wrap over an interval), and one kernel turns the sums into the interval
// mean mass and concentration per cell, so only the finished fields cross
// PCIe.  write_gridded_output_netcdf (io/netcdf_output.rs:141-229) becomes a
// host writer of the same dimensions, variables and attributes in the
// netCDF-3 64-bit-data format (CDF-5), which keeps the reference's u32 and
// i64 variable types, or in the classic 64-bit-offset format (CDF-2).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fp_internal.hpp"

namespace fp {
namespace {

constexpr uint32_t kOutBlock = 256;

__global__ void __launch_bounds__(kOutBlock) output_sample_kernel(const uint32_t *counts, const uint32_t *mass,
                                                                  unsigned long long *acc_count,
                                                                  unsigned long long *acc_mass, uint64_t n)
{
    const uint64_t c = (uint64_t)blockIdx.x * kOutBlock + threadIdx.x;
    if (c >= n) return;
    acc_count[c] += counts[c];
    acc_mass[c] += mass[c];
}

// mass_kg = (f32)(sum / n) / mass_scale: for one sample exactly the
// reference's (value as f32) / mass_scale, since a u32 converts to binary64
// exactly; conc = mass_kg / V (vol: nz floats per output row iy).
__global__ void __launch_bounds__(kOutBlock) output_finalize_kernel(const unsigned long long *acc_count,
                                                                    const unsigned long long *acc_mass,
                                                                    uint32_t n_samples, float mass_scale,
                                                                    const float *vol, uint32_t ny, uint32_t nz,
                                                                    uint32_t *count_out, float *mass_out,
                                                                    float *conc_out, uint64_t n)
{
    const uint64_t c = (uint64_t)blockIdx.x * kOutBlock + threadIdx.x;
    if (c >= n) return;
    const unsigned long long ac = acc_count[c];
    count_out[c] = ac > 0xffffffffull ? 0xffffffffu : (uint32_t)ac;
    const double mean = (double)acc_mass[c] / (double)n_samples;
    const float m = (float)mean / mass_scale;
    mass_out[c] = m;
    if (conc_out) {
        const uint32_t k = (uint32_t)(c % nz), iy = (uint32_t)((c / nz) % ny);
        conc_out[c] = m / vol[(uint64_t)iy * nz + k];
    }
}

}  // namespace

hipError_t launch_output_sample(const uint32_t *counts, const uint32_t *mass, unsigned long long *acc_count,
                                unsigned long long *acc_mass, uint64_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    output_sample_kernel<<<(uint32_t)((n + kOutBlock - 1) / kOutBlock), kOutBlock, 0, st>>>(counts, mass, acc_count,
                                                                                            acc_mass, n);
    return hipGetLastError();
}

hipError_t launch_output_finalize(const unsigned long long *acc_count, const unsigned long long *acc_mass,
                                  uint32_t n_samples, float mass_scale, const float *vol, uint32_t ny, uint32_t nz,
                                  uint32_t *count_out, float *mass_out, float *conc_out, uint64_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    output_finalize_kernel<<<(uint32_t)((n + kOutBlock - 1) / kOutBlock), kOutBlock, 0, st>>>(
        acc_count, acc_mass, n_samples, mass_scale, vol, ny, nz, count_out, mass_out, conc_out, n);
    return hipGetLastError();
}

// Cell volumes V = (dx * dx_m(lat_c)) * (dy * dy_m) * dz_k [m^3] per output
// row iy and level k: dx_m(lat) = R cos(lat PI180) PI180 is
// dx_meters_at_latitude (coords/mod.rs:156-158) in f32, dy_m = R PI180 the
// same at the equator, lat_c the row's centre latitude, dz_k the layer
// between outheights[k-1] (0 below the first) and outheights[k].
void output_volumes(const fp_output_geometry &g, const float *outheights, uint32_t ny, uint32_t nz, float *vol)
{
    const float r_earth = 6.371e6f;                          // lib.rs:22
    const float pi180 = 3.14159265358979323846f / 180.0f;    // lib.rs:32
    const double dy_m = (double)(r_earth * pi180);
    for (uint32_t iy = 0; iy < ny; ++iy) {
        const float lat = (float)(g.ylat0 + ((double)iy + 0.5) * g.dy);
        const double dx_m = (double)(r_earth * std::cos(lat * pi180) * pi180);
        const double area = (g.dx * dx_m) * (g.dy * dy_m);
        for (uint32_t k = 0; k < nz; ++k) {
            const double dz = (double)outheights[k] - (k ? (double)outheights[k - 1] : 0.0);
            vol[(uint64_t)iy * nz + k] = (float)(area * dz);
        }
    }
}

// ------------------------------------------------------------------ netCDF

namespace {

enum NcType : uint32_t { NC_CHAR = 2, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6, NC_UINT = 9, NC_INT64 = 10 };

uint32_t nc_size(uint32_t t) { return t == NC_CHAR ? 1u : (t == NC_DOUBLE || t == NC_INT64 ? 8u : 4u); }

struct NcAtt {
    std::string name;
    uint32_t type;
    std::vector<uint8_t> be;  // big-endian values
    uint64_t n;
};
struct NcVar {
    std::string name;
    std::vector<uint32_t> dims;
    std::vector<NcAtt> atts;
    uint32_t type;
    std::vector<uint8_t> be;  // big-endian data
};

class Be {  // big-endian byte writer with netCDF padding
public:
    std::vector<uint8_t> b;
    void u32(uint32_t v) { for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s)); }
    void u64(uint64_t v) { for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s)); }
    void raw(const void *p, size_t n) { b.insert(b.end(), (const uint8_t *)p, (const uint8_t *)p + n); }
    void pad() { while (b.size() % 4) b.push_back(0); }
};

void be_values(std::vector<uint8_t> &out, const void *src, uint64_t n, uint32_t size)
{
    const uint8_t *s = (const uint8_t *)src;
    out.resize(n * size);
    for (uint64_t i = 0; i < n; ++i)
        for (uint32_t k = 0; k < size; ++k) out[i * size + k] = s[i * size + (size - 1 - k)];  // little -> big endian
}

NcAtt att_text(const std::string &name, const std::string &v)
{
    NcAtt a{name, NC_CHAR, std::vector<uint8_t>(v.begin(), v.end()), v.size()};
    return a;
}

NcAtt att_int(const std::string &name, int32_t v)
{
    NcAtt a{name, NC_INT, {}, 1};
    be_values(a.be, &v, 1, 4);
    return a;
}

// netCDF-3 file: header (dimensions, global attributes, variables with their
// begin offsets) followed by the non-record variables' data in order.
// version 5: NON_NEG fields are 64-bit (CDF-5); version 2: 32-bit (CDF-2);
// OFFSET is 64-bit in both.
std::vector<uint8_t> encode(int version, const std::vector<std::pair<std::string, uint64_t>> &dims,
                            const std::vector<NcAtt> &gatts, std::vector<NcVar> &vars)
{
    auto build = [&](const std::vector<uint64_t> &begins) {
        Be h;
        h.raw("CDF", 3);
        h.b.push_back((uint8_t)version);
        auto nonneg = [&](uint64_t v) { if (version == 5) h.u64(v); else h.u32((uint32_t)v); };
        auto name = [&](const std::string &s) { nonneg(s.size()); h.raw(s.data(), s.size()); h.pad(); };
        auto atts = [&](const std::vector<NcAtt> &as) {
            if (as.empty()) { h.u32(0); nonneg(0); return; }
            h.u32(0x0C);
            nonneg(as.size());
            for (const NcAtt &a : as) {
                name(a.name);
                h.u32(a.type);
                nonneg(a.n);
                h.raw(a.be.data(), a.be.size());
                h.pad();
            }
        };
        nonneg(0);  // numrecs: no record dimension
        h.u32(0x0A);
        nonneg(dims.size());
        for (const auto &d : dims) { name(d.first); nonneg(d.second); }
        atts(gatts);
        h.u32(0x0B);
        nonneg(vars.size());
        for (size_t i = 0; i < vars.size(); ++i) {
            const NcVar &v = vars[i];
            name(v.name);
            nonneg(v.dims.size());
            for (uint32_t d : v.dims) nonneg(d);
            atts(v.atts);
            h.u32(v.type);
            nonneg((v.be.size() + 3) / 4 * 4);  // vsize
            h.u64(begins[i]);                   // OFFSET
        }
        return h.b;
    };
    std::vector<uint64_t> begins(vars.size(), 0);
    const uint64_t header = build(begins).size();  // the header size does not depend on the offsets
    uint64_t off = header;
    for (size_t i = 0; i < vars.size(); ++i) {
        begins[i] = off;
        off += (vars[i].be.size() + 3) / 4 * 4;
    }
    std::vector<uint8_t> file = build(begins);
    for (const NcVar &v : vars) {
        file.insert(file.end(), v.be.begin(), v.be.end());
        while (file.size() % 4) file.push_back(0);
    }
    return file;
}

}  // namespace

// write_gridded_output_netcdf (io/netcdf_output.rs:141-229).
int write_netcdf(const char *path, const fp_netcdf_snapshot &s, int version, std::string &err)
{
    const uint64_t nxyz = (uint64_t)s.nx * s.ny * s.nz, nxy = (uint64_t)s.nx * s.ny;
    std::vector<NcAtt> g = {att_text("title", "FLEXPART-GPU concentration/deposition MVP output"),
                            att_text("source", "flexpart-gpu (I-03 NetCDF concentration/deposition writer)"),
                            att_text("conventions", "CF-1.10"), att_text("deterministic_fields", "true"),
                            att_int("species_index", s.species_index)};
    std::vector<NcVar> vars;
    {  // time: i64 seconds since the epoch (CDF-2 has no 64-bit integer: binary64)
        NcVar v{"time", {0}, {att_text("units", "seconds since 1970-01-01 00:00:00 UTC")}, NC_INT64, {}};
        if (version == 5) {
            be_values(v.be, &s.time_s, 1, 8);
        } else {
            const double t = (double)s.time_s;
            v.type = NC_DOUBLE;
            be_values(v.be, &t, 1, 8);
        }
        vars.push_back(v);
    }
    const uint32_t ns[3] = {s.nx, s.ny, s.nz};
    const char *axes[3] = {"X", "Y", "Z"};
    const char *names[3] = {"x", "y", "z"};
    for (int a = 0; a < 3; ++a) {
        std::vector<float> idx(ns[a]);
        for (uint32_t k = 0; k < ns[a]; ++k) idx[k] = (float)k;
        NcVar v{names[a], {(uint32_t)(a + 1)}, {att_text("axis", axes[a])}, NC_FLOAT, {}};
        be_values(v.be, idx.data(), ns[a], 4);
        vars.push_back(v);
    }
    {
        NcVar v{"concentration_particle_count", {0, 1, 2, 3},
                {att_text("long_name", "particle count per concentration cell"), att_text("units", "1")}, NC_UINT, {}};
        if (version != 5) {  // CDF-2: i32 (counts above 2^31 - 1 are rejected)
            for (uint64_t i = 0; i < nxyz; ++i)
                if (s.count[i] > 0x7fffffffu) { err = "count above 2^31-1 needs the CDF-5 format"; return FP_E_INVALID; }
            v.type = NC_INT;
        }
        be_values(v.be, s.count, nxyz, 4);
        vars.push_back(v);
    }
    {
        NcVar v{"concentration_mass_kg", {0, 1, 2, 3},
                {att_text("long_name", "accumulated species mass per concentration cell"), att_text("units", "kg")},
                NC_FLOAT, {}};
        be_values(v.be, s.mass_kg, nxyz, 4);
        vars.push_back(v);
    }
    const float *dep[2] = {s.dry_kg_m2, s.wet_kg_m2};
    const char *dep_name[2] = {"dry_deposition_kg_m2", "wet_deposition_kg_m2"};
    const char *dep_long[2] = {"dry deposition per horizontal cell", "wet deposition per horizontal cell"};
    std::vector<float> zeros;
    for (int d = 0; d < 2; ++d) {
        const float *src = dep[d];
        if (!src) {
            zeros.assign(nxy, 0.0f);
            src = zeros.data();
        }
        NcVar v{dep_name[d], {0, 1, 2}, {att_text("long_name", dep_long[d]), att_text("units", "kg m-2")}, NC_FLOAT, {}};
        be_values(v.be, src, nxy, 4);
        vars.push_back(v);
    }
    if (s.conc_kg_m3) {  // addition: the volume-normalised concentration
        NcVar v{"concentration_kg_m3", {0, 1, 2, 3},
                {att_text("long_name", "mean concentration per cell (mass / (mass_scale * cell volume))"),
                 att_text("units", "kg m-3")},
                NC_FLOAT, {}};
        be_values(v.be, s.conc_kg_m3, nxyz, 4);
        vars.push_back(v);
    }
    const std::vector<std::pair<std::string, uint64_t>> dims = {{"time", 1}, {"x", s.nx}, {"y", s.ny}, {"z", s.nz}};
    const std::vector<uint8_t> file = encode(version, dims, g, vars);
    FILE *f = std::fopen(path, "wb");
    if (!f) { err = std::string("cannot open ") + path; return FP_E_INVALID; }
    const size_t w = std::fwrite(file.data(), 1, file.size(), f);
    const int closed = std::fclose(f);
    if (w != file.size() || closed != 0) { err = std::string("write failed: ") + path; return FP_E_INVALID; }
    return FP_OK;
}

}  // namespace fp
