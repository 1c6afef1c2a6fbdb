/*
 * oracle.h - CPU restatement of FLEXPART-GPU's production per-particle
 * timestep.  TEST INFRASTRUCTURE ONLY: this library is the checker for the
 * HIP path.  Only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark load it; the product never does.
 *
 * State is the reference's 96-byte AoS particle (particles/mod.rs:76-168),
 * met fields use the reference flat layout ((ix*ny)+iy)*nz+iz
 * (advection_dual_wind.wgsl:82-84) and PBL fields i*ny+j
 * (langevin_fused.wgsl:209-211).
 */
#ifndef FP_ORACLE_H
#define FP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float pos_x, pos_y, pos_z;
    int32_t cell_x, cell_y;
    uint32_t flags;
    float mass[4];
    float vel_u, vel_v, vel_w;
    float turb_u, turb_v, turb_w;
    int32_t time, timestep, time_mem, time_split, release_point, class_id, cbt;
    uint32_t pad0;
} orc_particle;

typedef struct {
    uint32_t nx, ny, nz;
    const float *u0, *v0, *w0, *u1, *v1, *w1;
    const float *level_heights; /* nz entries; top <= 0 selects level-index z */
} orc_met;

typedef struct {
    uint32_t nx, ny;
    const float *ustar, *wstar, *hmix, *oli;
    const float *precip; /* mm/h, used by wet_mode 3 only */
} orc_pbl;

typedef struct {
    float dt_seconds;
    float alpha;
    float x_scale, y_scale, z_scale;
    int32_t direction; /* +1 forward, -1 backward (advection only) */
    uint32_t key[2];
    uint32_t counter[4];
    uint32_t n_substeps;
    float min_height_m;
    float rho_grad_over_rho;
    uint32_t dry_mode; /* 0 off, 1 scalar v_d, 2 per-slot v_d */
    float dry_vd;
    float dry_href;
    const float *dry_vd_slots;
    uint32_t wet_mode; /* 0 off, 1 scalar, 2 per-slot, 3 lambda = A * P^B */
    float wet_lambda, wet_fraction;
    const float *wet_lambda_slots, *wet_fraction_slots;
    float wet_a, wet_b;
    uint32_t loss_mode; /* 0 none, 1 deactivate when mass[0] < loss_min_mass */
    float loss_min_mass;
    int32_t n_threads;
} orc_step_params;

typedef struct {
    uint32_t nx, ny, nz;
    uint32_t refine_x, refine_y; /* output cells per met cell (1 = met grid) */
    uint32_t species;
    float mass_scale;
    float outheights[16];
} orc_grid_params;

void orc_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]);
void orc_counter_add(const uint32_t counter[4], uint64_t offset, uint32_t out[4]);
void orc_philox_stream(const uint32_t counter[4], const uint32_t key[2], uint64_t n_blocks,
                       uint32_t *out);
float orc_u32_to_uniform01(uint32_t v);
void orc_box_muller(float u0, float u1, float out[2]);

int orc_step(orc_particle *p, uint64_t n, const orc_met *met, const orc_pbl *pbl,
             const orc_step_params *sp, uint32_t next_counter[4]);

float orc_update_horizontal(float prev, float sigma, float tl, float dt, float eta);
float orc_update_vertical(float prev, float sigw, float dsigwdz, float tlw, float dt, float rho,
                          float eta);
void orc_deposit(orc_particle *p, uint64_t n, const orc_pbl *pbl, const orc_step_params *sp);
void orc_advect(orc_particle *p, uint64_t n, const orc_met *met, const orc_step_params *sp);

void orc_hanna(float ust, float wst, float ol, float h, float z, float out[13]);
/* PBL reflection of one particle (langevin_fused.wgsl:481-500), in place */
void orc_reflect(float *z, float *turb_w, int32_t *cbt, float z_min, float h);
/* hanna_short_w (langevin_fused.wgsl:296-344): sigma_w, dsigma_w/dz, T_Lw */
void orc_hanna_short_w(float ust, float wst, float ol, float h, float z, float out[3]);

void orc_grid_accumulate(const orc_particle *p, uint64_t n, const orc_grid_params *gp,
                         uint32_t *counts, uint32_t *mass_fp);

uint64_t orc_compact(const orc_particle *p, uint64_t n, orc_particle *out,
                     uint32_t *survivor_map);

uint64_t orc_morton_encode_3d_21bit(uint32_t x, uint32_t y, uint32_t z);
void orc_morton_keys(const orc_particle *p, uint64_t n, uint32_t bits, int explicit_bounds,
                     const double bounds[6], uint64_t *keys);
int orc_sort_morton(const orc_particle *p, uint64_t n, uint32_t bits, int explicit_bounds,
                    const double bounds[6], orc_particle *out, uint32_t *new_to_old);

void orc_math_eval(int fn, const float *x, const float *y, float *out, uint64_t n);
/* meters_to_grid_level (advection_dual_wind.wgsl:94-117) of n heights over
 * the nz-level table (level-index mode when its top is not positive). */
void orc_level_eval(const float *level_heights, uint32_t nz, const float *z, float *out, uint64_t n);

/* ---- source-term injection (release/mod.rs) --------------------------- */

typedef struct {
    double xlon0, ylat0, dx, dy; /* GridDomain (coords/mod.rs:72-85) */
    uint32_t nx, ny;
} orc_grid_domain;

typedef struct {
    double lon_min, lon_max, lat_min, lat_max, z_min, z_max; /* ReleaseBounds */
    double mass_kg;
    uint64_t particle_count;
    int64_t start_s, end_s;
    int32_t release_point, class_id;
} orc_release;

/* ReleaseManager::inject_for_time (release/mod.rs:90-152): for every release
 * in order, the particles between released[r] and the target count at now_s
 * (target_emitted_count, :181-197), placed by sample_stratified (:264-283)
 * and geo_to_grid / to_relative (coords/mod.rs:109-114, 169-178), written to
 * out[0..] in emission order.  released[] is advanced per release.  Returns
 * the number written; *status = 0, 1 (a sample fell outside the grid: the
 * particles before it are written and that release's count is not advanced)
 * or 2 (more than max_out particles: max_out are written). */
uint64_t orc_release_inject(const orc_grid_domain *grid, const orc_release *rel, uint32_t n_rel,
                            uint64_t *released, int64_t now_s, int32_t particle_time,
                            orc_particle *out, uint64_t max_out, int *status);

/* PBL options read by the GPU diagnostics (PblComputationOptions,
 * io/pbl_params.rs:36-71; pbl_diagnostics.wgsl:45-58). */
typedef struct {
    float roughness_length_m, wind_reference_height_m, heat_flux_neutral_threshold_w_m2;
    float fallback_mixing_height_m, hmix_min_m, hmix_max_m;
} orc_pbl_options;

/* Surface fields in the order of SurfaceCellInput (pbl_diagnostics.wgsl:30-43):
 * pressure, T2m, u10, v10, stress, sshf, ssr, hmix, ustar, 1/L. */
#define ORC_SURF_COUNT 10

/* interpolate_surface_fields_linear (io/temporal.rs:122-214) of the ten
 * fields with weight alpha, then the GPU PBL diagnostics per cell
 * (pbl_diagnostics.wgsl:220-305).  out[0..6]: ustar, wstar, hmix, oli, sshf,
 * ssr, surfstr (n floats each). */
void orc_pbl_from_surface(uint64_t n, const float *const s0[ORC_SURF_COUNT], const float *const s1[ORC_SURF_COUNT],
                          float alpha, const orc_pbl_options *o, float *const out[7]);

/* transform_hybrid_sigma_to_height (io/vertical_transform.rs:144-212) and
 * compute_hybrid_level_pressure_pa (:106-137), in the reference's loop order;
 * the mean profile is its sequential f32 sum (:315-334).  Returns 0 or the
 * error kind (1 non-positive pressure, 2 non-decreasing pressure, 3 invalid
 * thermodynamics, 4 non-monotonic height) with err_ijk = (i, j, k). */
int orc_hybrid_to_height(uint32_t nx, uint32_t ny, uint32_t nz, const float *ps, const float *t, const float *q,
                         const float *ak, const float *bk, float *pressure, float *height, float *mean,
                         uint32_t err_ijk[3]);

/* Concentration output.  Volumes of the output cells per row iy and level k
 * (docs/science/concentration-gridding.md:35-41 with dx_meters_at_latitude,
 * coords/mod.rs:156-158): vol[iy*nz+k]. */
void orc_output_volumes(double ylat0, double dx, double dy, const float *outheights, uint32_t ny, uint32_t nz,
                        float *vol);
/* Interval fields from 64-bit sums: count (saturating), mean mass [kg] =
 * (f32)(sum/n)/mass_scale (gpu/gridding.rs:210-213 for n = 1) and, when vol
 * is not NULL, mass / vol. */
void orc_output_finalize(uint64_t ncell, const uint64_t *acc_count, const uint64_t *acc_mass, uint32_t n_samples,
                         float mass_scale, const float *vol, uint32_t ny, uint32_t nz, uint32_t *count, float *mass,
                         float *conc);

#ifdef __cplusplus
}
#endif

#endif
