/*
 * flexpart_hip.h - C ABI of the MI355X (gfx950) per-particle timestep.
 *
 * This library replaces the reference's GPU backend (src/gpu + src/shaders of
 * the flexpart-gpu crate) behind plain C calls: the crate's time loop
 * (simulation/timeloop.rs) keeps its control flow and calls these entry
 * points where it encodes wgpu passes today.  Every entry point names the
 * reference interface it replaces.
 *
 * Conventions
 *   - Status: 0 = OK, negative = error class (FP_E_*); fp_last_error() holds
 *     the message.  Nothing unwinds across the ABI.
 *   - A context owns one HIP stream and every device buffer (RAII in the
 *     reference: the driver struct owns every wgpu buffer, timeloop.rs:713-828).
 *     Calls on one context are stream-ordered and not thread-safe.
 *     fp_sync() and calls that return host data block; the rest are async.
 *   - Host pointers are plain host memory; "_dev" entry points take device
 *     pointers already resident in HBM.
 *   - Particle state is stored as SoA in HBM.  The AoS exchange format is the
 *     reference's 96-byte Particle (particles/mod.rs:76-168).
 */
#ifndef FLEXPART_HIP_H
#define FLEXPART_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FP_OK 0
#define FP_E_INVALID (-1)   /* bad argument / shape */
#define FP_E_DEVICE (-2)    /* HIP runtime error */
#define FP_E_NOMEM (-3)     /* device allocation failed */
#define FP_E_STATE (-4)     /* call order (e.g. step before met upload) */

#define FP_FLAG_ACTIVE 1u
#define FP_MAX_SPECIES 4
#define FP_MAX_OUTPUT_LEVELS 16

/* Reference Particle, 96 bytes, repr(C) (particles/mod.rs:76-168). */
typedef struct fp_particle {
    float pos_x, pos_y, pos_z;
    int32_t cell_x, cell_y;
    uint32_t flags;
    float mass[FP_MAX_SPECIES];
    float vel_u, vel_v, vel_w;
    float turb_u, turb_v, turb_w;
    int32_t time, timestep, time_mem, time_split, release_point, class_id, cbt;
    uint32_t pad0;
} fp_particle;

typedef struct fp_ctx fp_ctx;

/* Context shape.  met_* is the wind grid (WindField3D shape, wind/mod.rs:138),
 * the PBL grid is met_nx x met_ny (PblState, pbl/mod.rs:130-174). */
typedef struct fp_ctx_desc {
    uint64_t capacity;      /* particle slots (ParticleStore::with_capacity), at most 2^30 */
    uint32_t met_nx, met_ny, met_nz; /* met_nx*met_ny*met_nz <= 2^27 (a 4 GiB packed bracket) */
    uint32_t n_species;     /* mass slots the step updates, 1..4 */
    uint32_t out_nx, out_ny, out_nz; /* concentration grid (0 = none) */
    uint32_t reserved[6];
} fp_ctx_desc;

/* Per-step parameters.  Gathers DualWindAdvectionParams (gpu/advection.rs:
 * 58-73), FusedLangevinParams (gpu/langevin_fused.rs:98-115) and the
 * deposition dispatch params (gpu/deposition.rs, gpu/wet_deposition.rs). */
typedef struct fp_step_params {
    float dt_seconds;        /* > 0 */
    float alpha;             /* temporal weight, clamped to [0,1] */
    float x_scale, y_scale, z_scale; /* VelocityToGridScale */
    int32_t direction;       /* +1 forward, -1 backward (advection dt sign) */
    uint32_t key[2];         /* Philox key */
    uint32_t counter[4];     /* Philox base counter for this step */
    uint32_t n_substeps;     /* vertical sub-steps 0..4 (timeloop.rs:1111: 4) */
    float min_height_m;      /* reflection floor (timeloop.rs:1112: 0.01) */
    float rho_grad_over_rho;
    uint32_t dry_mode;       /* FP_DRY_* */
    float dry_vd;            /* m/s, FP_DRY_UNIFORM */
    float dry_href;          /* reference height (constants::HREF = 15 m) */
    uint32_t wet_mode;       /* FP_WET_* */
    float wet_lambda;        /* 1/s, FP_WET_UNIFORM */
    float wet_fraction;      /* FP_WET_UNIFORM and FP_WET_PRECIP_FIELD */
    float wet_a, wet_b;      /* lambda = a * P^b, FP_WET_PRECIP_FIELD */
    uint32_t loss_mode;      /* FP_LOSS_* */
    float loss_min_mass;     /* kg, FP_LOSS_MIN_MASS */
    uint32_t grid_fused;     /* 1: accumulate the concentration grid after the step */
    uint32_t reserved[7];
} fp_step_params;

#define FP_DRY_OFF 0u
#define FP_DRY_UNIFORM 1u
#define FP_DRY_PER_SLOT 2u   /* v_d[slot] from fp_forcing_upload */
#define FP_WET_OFF 0u
#define FP_WET_UNIFORM 1u
#define FP_WET_PER_SLOT 2u   /* lambda[slot], fraction[slot] */
#define FP_WET_PRECIP_FIELD 3u /* lambda = a * precip(i,j)^b, nearest PBL cell */
#define FP_LOSS_NONE 0u
#define FP_LOSS_MIN_MASS 1u  /* clear FLAG_ACTIVE when mass[0] < loss_min_mass */

/* Concentration gridding (ConcentrationGriddingParams, gpu/gridding.rs:90-110). */
typedef struct fp_grid_params {
    uint32_t refine_x, refine_y; /* output cells per met cell; 1 = met grid */
    uint32_t species;
    float mass_scale;            /* fixed-point scale (reference default 1e6) */
    float outheights[FP_MAX_OUTPUT_LEVELS]; /* metres; all zero = level index */
} fp_grid_params;

/* Morton re-sort (ParticleSpatialSortOptions, particles/mod.rs:321-340). */
typedef struct fp_sort_params {
    uint32_t bits_per_axis;      /* 1..21 */
    uint32_t explicit_bounds;    /* 0: active-particle extent */
    double bounds[6];            /* x_min,x_max,y_min,y_max,z_min,z_max */
} fp_sort_params;

/* ---- context -------------------------------------------------------- */

/* Replaces GpuContext::new (gpu/mod.rs:163-208) + ParticleBuffers::from_store
 * (gpu/buffers.rs:381-420): selects the device, allocates every buffer. */
int fp_ctx_create(int device_ordinal, const fp_ctx_desc *desc, fp_ctx **out);
int fp_ctx_destroy(fp_ctx *ctx);
/* Last error message of this context (or of the last failed create). */
const char *fp_last_error(const fp_ctx *ctx);
/* Blocks until all work on the context stream is done
 * (device.poll(Maintain::Wait), timeloop.rs:1051-1058). */
int fp_sync(fp_ctx *ctx);
/* Library version string. */
const char *fp_version(void);

/* ---- particles ------------------------------------------------------ */

/* ParticleBuffers::upload_particle_slots (gpu/buffers.rs:464-493). */
int fp_particles_upload_slots(fp_ctx *ctx, const uint32_t *slots, const fp_particle *p,
                              uint64_t n);
/* ParticleBuffers::upload_store (gpu/buffers.rs:440-462): slots [0, n). */
int fp_particles_upload(fp_ctx *ctx, const fp_particle *p, uint64_t n);
/* ParticleBuffers::download_particles (gpu/buffers.rs:111-166): slots [0, n). */
int fp_particles_download(fp_ctx *ctx, fp_particle *out, uint64_t n);
/* ParticleBuffers::set_dispatch_count (gpu/buffers.rs:424-436). */
int fp_particles_set_dispatch_count(fp_ctx *ctx, uint64_t n);
uint64_t fp_particles_dispatch_count(const fp_ctx *ctx);
/* Device pointers of the SoA arrays (field order of fp_particle, mass as
 * FP_MAX_SPECIES arrays); valid until the next compaction or sort. */
int fp_particles_soa_pointers(fp_ctx *ctx, void **ptrs, uint32_t n_ptrs);

/* ---- met / PBL / forcing -------------------------------------------- */

/* DualWindBuffers::from_fields / upload_t0 / upload_t1 (gpu/buffers.rs:801-953):
 * six fields of met_nx*met_ny*met_nz floats, flat ((ix*ny)+iy)*nz+iz, and the
 * level heights [m] (nz entries, non-decreasing; all zero = level-index z,
 * VelocityToGridScale::IDENTITY). */
int fp_met_upload_bracket(fp_ctx *ctx, const float *u0, const float *v0, const float *w0,
                          const float *u1, const float *v1, const float *w1,
                          const float *level_heights);
int fp_met_upload_bracket_dev(fp_ctx *ctx, const float *u0, const float *v0, const float *w0,
                              const float *u1, const float *v1, const float *w1,
                              const float *level_heights_host);
/* PblBuffers::upload_state (gpu/pbl.rs): ustar, wstar, hmix, oli (1/L) on
 * met_nx x met_ny, index i*ny+j; precip [mm/h] may be NULL. */
int fp_pbl_upload(fp_ctx *ctx, const float *ustar, const float *wstar, const float *hmix,
                  const float *oli, const float *precip);
/* ---- met ingest on the device ---------------------------------------- */

/* Pinned host memory (hipHostMalloc) for met fields: H2D copies from it run
 * on the DMA engines without a host-side staging copy.  NULL on failure. */
void *fp_host_alloc(uint64_t bytes);
int fp_host_free(void *p);

/* Asynchronous bracket upload: the six fields of the next met bracket are
 * copied on the context's copy stream while fp_step keeps using the current
 * bracket; fp_met_commit_prefetch then orders the context's stream after the
 * copies and packs the new bracket (as fp_met_upload_bracket does) without
 * blocking the host.  The replacement for the reference's prefetch pipeline
 * (io/grib2_async.rs:49-96 decoding on a background thread, then the
 * DualWindBuffers uploads of timeloop.rs:1378-1402).  The host arrays must
 * stay untouched until the commit returns; use fp_host_alloc memory for a
 * transfer that overlaps the steps.  One prefetch may be pending at a time. */
int fp_met_prefetch_bracket(fp_ctx *ctx, const float *u0, const float *v0, const float *w0,
                            const float *u1, const float *v1, const float *w1,
                            const float *level_heights);
int fp_met_commit_prefetch(fp_ctx *ctx);

/* transform_hybrid_sigma_to_height (io/vertical_transform.rs:144-212) on the
 * met grid: p = ak + bk ps per level, virtual temperature, hypsometric layer
 * thicknesses into geometric heights per column, and the horizontal mean
 * height profile (the grid's level heights, for fp_met_upload_bracket).
 * ps: met_nx*met_ny; temperature, humidity (NULL = dry): met_nx*met_ny*met_nz,
 * index ((i*ny)+j)*nz+k; ak_pa, bk: met_nz.  mean_heights: met_nz floats
 * (host); pressure_out / height_out: optional host arrays of the per-column
 * fields.  A transform error returns FP_E_INVALID with *err (may be NULL)
 * holding the reference's error kind and its (i, j, k). */
#define FP_VT_OK 0u
#define FP_VT_NONPOSITIVE_PRESSURE 1u
#define FP_VT_NONDECREASING_PRESSURE 2u
#define FP_VT_INVALID_THERMODYNAMICS 3u
#define FP_VT_NONMONOTONIC_HEIGHT 4u
typedef struct fp_vt_error {
    uint32_t code;      /* FP_VT_* */
    uint32_t i, j, k;
} fp_vt_error;
int fp_met_hybrid_to_height(fp_ctx *ctx, const float *ps, const float *temperature, const float *humidity,
                            const float *ak_pa, const float *bk, float *mean_heights, float *pressure_out,
                            float *height_out, fp_vt_error *err);

/* ---- PBL from surface fields on the device ---------------------------- */

/* The ten surface fields the PBL diagnostics read, in the order of the
 * reference's packed SurfaceCellInput (shaders/pbl_diagnostics.wgsl:30-43;
 * SurfaceFields, wind/mod.rs:194-225): met_nx x met_ny floats each, index
 * i*ny+j. */
#define FP_SURF_PRESSURE 0      /* surface pressure [Pa] */
#define FP_SURF_T2M 1           /* 2 m temperature [K] */
#define FP_SURF_U10 2           /* 10 m wind [m/s] */
#define FP_SURF_V10 3
#define FP_SURF_STRESS 4        /* surface stress [N/m^2] */
#define FP_SURF_SSHF 5          /* sensible heat flux [W/m^2], positive upward */
#define FP_SURF_SSR 6           /* solar radiation [W/m^2] */
#define FP_SURF_HMIX 7          /* met mixing height [m] (<= 0: not provided) */
#define FP_SURF_USTAR 8         /* met friction velocity [m/s] (<= 0 / non-finite: not provided) */
#define FP_SURF_OLI 9           /* met 1/L [1/m] (|.| < 1e-5 / non-finite: not provided) */
#define FP_SURF_COUNT 10

/* PblComputationOptions (io/pbl_params.rs:36-71), the part the GPU
 * diagnostics read (pbl_diagnostics.wgsl:45-58).  Reference defaults: 0.1,
 * 10, 1, 800, 100, 4500. */
typedef struct fp_pbl_options {
    float roughness_length_m;
    float wind_reference_height_m;
    float heat_flux_neutral_threshold_w_m2;
    float fallback_mixing_height_m;
    float hmix_min_m;
    float hmix_max_m;
} fp_pbl_options;

/* The surface fields of met time level `level` (0: t0, 1: t1) valid at
 * time_s, kept on the device until replaced (met bracket changes):
 * SurfaceFieldBuffer::upload (gpu/pbl.rs:208-229) with both bracket ends
 * resident.  fields[k] is field FP_SURF_k; host pointers (_dev: device
 * pointers, copied on the context's stream). */
int fp_surface_upload(fp_ctx *ctx, int level, const float *const fields[FP_SURF_COUNT], int64_t time_s);
int fp_surface_upload_dev(fp_ctx *ctx, int level, const float *const fields[FP_SURF_COUNT], int64_t time_s);
/* One kernel per cell: interpolate_surface_fields_linear (io/temporal.rs:
 * 122-214; alpha as fp_interpolation_alpha with `strict`, exact copies at
 * alpha <= 0 and >= 1) followed by the PBL diagnostics of
 * pbl_diagnostics.wgsl:220-305 (timeloop.rs:961-993 with GPU PBL).  The
 * result replaces the context's PBL state as fp_pbl_upload would (precip is
 * left unchanged).  passthrough_dev (device, NULL to skip) receives
 * 3 x met_nx*met_ny floats: sanitized sshf, ssr and surface stress (the
 * reference's out_sshf / out_ssr / out_surfstr).  Asynchronous on the
 * context's stream (the next fp_step is ordered after it); fp_sync before
 * reading passthrough_dev from another stream. */
int fp_pbl_from_surface(fp_ctx *ctx, const fp_pbl_options *options, int64_t now_s, int strict,
                        float *passthrough_dev);
/* PblBuffers::download_state (gpu/pbl.rs): the context's current ustar,
 * wstar, hmix and 1/L (host arrays of met_nx*met_ny floats; any may be NULL). */
int fp_pbl_download(fp_ctx *ctx, float *ustar, float *wstar, float *hmix, float *oli);

/* DryDepositionIoBuffers / WetDepositionIoBuffers uploads: per-slot forcing
 * (capacity floats each; NULL leaves that array unchanged). */
int fp_forcing_upload(fp_ctx *ctx, const float *dry_vd, const float *wet_lambda,
                      const float *wet_fraction);

/* ---- the step ------------------------------------------------------- */

/* One production timestep over the dispatch range: the four reference passes
 * encode_advection_dual_wind_gpu_with_kernel (gpu/advection.rs:605),
 * encode_langevin_fused_gpu (gpu/langevin_fused.rs:133),
 * encode_dry_deposition_probability_gpu_with_kernel (gpu/deposition.rs:315),
 * encode_wet_deposition_probability_gpu_with_kernel (gpu/wet_deposition.rs:368)
 * fused in one kernel.  next_counter = counter + dispatch*(n_sub<=2 ? 1 : 2)
 * (gpu/langevin_fused.rs:229-233). */
int fp_step(fp_ctx *ctx, const fp_step_params *params, uint32_t next_counter[4]);

/* Vector-Jacobian product of one fp_step at the context's current particles
 * (the step is not applied): for every dispatched slot, lam_in = J^T lam_out
 * where J is the derivative of the step's output (x, y [grid units], z [m],
 * turb_u, turb_v, turb_w, mass[0]) with respect to the same quantities at
 * its input, under the step's own Philox draws.  lam_out / lam_in: 7 x
 * dispatch floats (host, quantity-major).  wind_u / wind_v (host, met_nx *
 * met_ny doubles, both or neither) receive the derivative of sum(lam_out *
 * output) with respect to a horizontal-wind offset at each met column (added
 * to every level and time level).  Inactive slots pass lam through.  The
 * building block of fp_adj_backward; no counterpart in the reference.
 * Blocking. */
int fp_step_vjp(fp_ctx *ctx, const fp_step_params *params, const float *lam_out, float *lam_in,
                double *wind_u, double *wind_v);

/* ---- concentration grid --------------------------------------------- */

/* ConcentrationGridBuffers::clear (gpu/gridding.rs:178-188). */
int fp_grid_clear(fp_ctx *ctx);
/* Sets the gridding parameters used by fp_grid_accumulate and by fused
 * gridding in fp_step. */
int fp_grid_configure(fp_ctx *ctx, const fp_grid_params *params);
/* dispatch_concentration_gridding_gpu (gpu/gridding.rs:255-407). */
int fp_grid_accumulate(fp_ctx *ctx);
/* download_output (gpu/gridding.rs:191-218), raw u32 sums (blocking). */
int fp_grid_download(fp_ctx *ctx, uint32_t *counts, uint32_t *mass_fp);

/* ---- compaction and Morton re-sort ---------------------------------- */

/* encode_compaction_with_reorder (gpu/compaction.rs:447-505) +
 * download_active_count (:290) + set_dispatch_count: survivors keep ascending
 * slot order, the tail holds deactivated copies.  survivor_map (host, may be
 * NULL) receives new_slot -> old_slot for the active prefix.  Blocking. */
int fp_compact(fp_ctx *ctx, uint64_t *active_count, uint32_t *survivor_map);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory - code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc: the survivor map and active count that fp_compact computes,
 * over a caller-owned device flag array, with no context and no allocation.
 * flags: n device words (bit 0 = FP_FLAG_ACTIVE); survivor_map: device, room
 * for n words, receives the active slots in ascending order; active_count:
 * one device word.  The kernels' scratch (one count per 2048 slots) is a
 * library-owned 256 KiB static device array per device, so n is at most
 * FP_SURVIVOR_MAP_STATIC_MAX (FP_E_INVALID beyond).  device_ordinal must be
 * the current device; stream: a hipStream_t (NULL = default stream).
 * Asynchronous.  Calls that overlap on one device share the scratch and race. */
#define FP_SURVIVOR_MAP_STATIC_MAX 134217728ull
int fp_survivor_map_static(int device_ordinal, const uint32_t *flags, uint64_t n, uint32_t *survivor_map,
                           uint32_t *active_count, void *stream);

/* ParticleStore::sort_spatially (particles/mod.rs:649-699) on the device:
 * stable sort of the dispatch range by Morton key, inactive slots last.
 * new_to_old (host, may be NULL) receives the permutation.  Blocking only
 * when new_to_old is given. */
int fp_sort_morton(fp_ctx *ctx, const fp_sort_params *params, uint32_t *new_to_old);

/* ---- source-term injection ------------------------------------------ */

/* GridDomain (coords/mod.rs:72-85): lon/lat of the lower-left corner and the
 * spacing, in degrees. */
typedef struct fp_grid_domain {
    double xlon0, ylat0, dx, dy;
    uint32_t nx, ny;
} fp_grid_domain;

/* One release (ReleaseConfig + ReleaseBounds, release/mod.rs:172-199): box
 * bounds, total mass and particle count, emission window [start_s, end_s]
 * in simulation seconds, and the release index written to its particles. */
typedef struct fp_release {
    double lon_min, lon_max, lat_min, lat_max, z_min, z_max;
    double mass_kg;
    uint64_t particle_count;
    int64_t start_s, end_s;
    int32_t release_point, class_id;
} fp_release;

/* 88 bytes */

/* ReleaseManager::inject_for_time (release/mod.rs:90-152) on the device: for
 * each release in order, the particles its window emits between released[r]
 * (in/out: emitted so far) and now_s are placed by stratified sampling and
 * written straight into the current particle set, in emission order, to
 * slots[0..n_slots) (the free-slot order of ParticleStore::add) or, with
 * slots == NULL, to first_slot, first_slot + 1, ... (the free slots after
 * ParticleStore::reset_after_compaction).  *injected receives the count.
 * A sample outside the grid domain or running out of slots writes the
 * particles before it, keeps that release's count and returns FP_E_INVALID,
 * as the reference does.  Blocking. */
int fp_release_inject(fp_ctx *ctx, const fp_grid_domain *grid, const fp_release *releases,
                      uint32_t n_releases, uint64_t *released, int64_t now_s,
                      const uint32_t *slots, uint64_t n_slots, uint64_t first_slot,
                      uint64_t *injected);

/* ---- concentration output ------------------------------------------- */

/* Output grid geometry for the cell volumes (OUTGRID: OUTLON0, OUTLAT0,
 * DXOUT, DYOUT in degrees; docs/science/concentration-gridding.md:9-20). */
typedef struct fp_output_geometry {
    double xlon0, ylat0;   /* south-west corner of the output grid [deg] */
    double dx, dy;         /* output cell size [deg] */
} fp_output_geometry;

/* Output intervals: fp_output_sample adds the current concentration grid
 * (fp_grid_accumulate or a grid_fused step) to per-cell 64-bit sums on the
 * device; fp_output_finalize turns the sums into the interval fields.  The
 * reference downloads and converts one instantaneous grid per output time
 * (accumulate_concentration_grid_gpu, gpu/gridding.rs:194-219); one sample
 * per interval reproduces that exactly. */
int fp_output_reset(fp_ctx *ctx);
int fp_output_sample(fp_ctx *ctx);
/* count_out: particle samples per cell over the interval (u32, saturating);
 * mass_out: interval-mean mass [kg] = (f32)(sum / n_samples) / mass_scale;
 * conc_out (NULL to skip; needs geometry and metre output heights): mass_out
 * / V_cell [kg m-3], V_cell = (dx * dx_meters_at_latitude(lat_c)) *
 * (dy * R PI180) * dz (docs/science/concentration-gridding.md:35-41,
 * coords/mod.rs:156-158).  Host arrays of out_nx*out_ny*out_nz, index
 * (ix*ny+iy)*nz+iz; n_samples may be NULL.  The sums are kept. */
int fp_output_finalize(fp_ctx *ctx, const fp_output_geometry *geometry, uint32_t *count_out, float *mass_out,
                       float *conc_out, uint32_t *n_samples);

/* GriddedOutputSnapshot (io/netcdf_output.rs:43-57) plus the optional
 * volume-normalised concentration. */
typedef struct fp_netcdf_snapshot {
    int64_t time_s;                /* simulation time [s since the epoch] */
    int32_t species_index;
    uint32_t nx, ny, nz;
    const uint32_t *count;         /* nx*ny*nz particle counts */
    const float *mass_kg;          /* nx*ny*nz */
    const float *dry_kg_m2;        /* nx*ny (NULL: zeros) */
    const float *wet_kg_m2;        /* nx*ny (NULL: zeros) */
    const float *conc_kg_m3;       /* nx*ny*nz or NULL (extra variable concentration_kg_m3) */
} fp_netcdf_snapshot;

/* write_gridded_output_netcdf (io/netcdf_output.rs:141-229): the same
 * dimensions (time = 1, x, y, z), variables, types and attributes.
 * format 5: netCDF-3 64-bit data (CDF-5; u32 counts, i64 time); format 2:
 * classic 64-bit offset (CDF-2; counts as i32, time as f64) for readers
 * without CDF-5.  Host only; errors via fp_last_error(NULL). */
int fp_write_netcdf(const char *path, const fp_netcdf_snapshot *snapshot, int format);

/* ---- temporal weight ------------------------------------------------ */

/* interpolation_alpha (simulation/timeloop.rs:2086-2120): the weight of the
 * later met level at target_s inside the bracket [t0_s, t1_s] (integer
 * seconds), (target - t0) / (t1 - t0) in binary64 rounded once to float.
 * A target outside the bracket is clamped (strict = 0) or rejected with
 * FP_E_INVALID (strict = 1); t1_s <= t0_s is FP_E_INVALID.  Host only. */
#define FP_TIME_CLAMP 0
#define FP_TIME_STRICT 1
int fp_interpolation_alpha(int64_t t0_s, int64_t t1_s, int64_t target_s, int strict,
                           float *alpha);

/* ---- RNG known-answer path ------------------------------------------ */

/* philox_rng.wgsl (gpu/rng.rs): n consecutive Philox4x32-10 blocks at
 * counter + i, written to out[4*i..4*i+3] (host).  Blocking. */
int fp_philox_blocks(fp_ctx *ctx, const uint32_t counter[4], const uint32_t key[2],
                     uint64_t n, uint32_t *out);

/* Evaluates one of the step kernel's arithmetic primitives on the device for
 * n inputs (host arrays; y is used by pow and div).  Diagnostic: lets a host
 * check that the device rounds exactly as the CPU does.  No counterpart in
 * the reference.  fn: FP_PROBE_*; FP_PROBE_LEVEL is the step's height ->
 * fractional level lookup (meters_to_grid_level) of the heights x over the
 * level table of the context's met bracket (FP_E_STATE before the first met
 * upload).  Blocking. */
#define FP_PROBE_EXP 0
#define FP_PROBE_LOG 1
#define FP_PROBE_SIN 2
#define FP_PROBE_COS 3
#define FP_PROBE_POW 4
#define FP_PROBE_SQRT 5
#define FP_PROBE_DIV 6
#define FP_PROBE_LOG_UNIT 7
#define FP_PROBE_LEVEL 8
int fp_math_probe(fp_ctx *ctx, int fn, const float *x, const float *y, float *out, uint64_t n);

/* Exhaustive device check of one of the step's arithmetic primitives over the
 * float bit patterns [first_bits, last_bits]; *failures receives the number of
 * inputs whose result is not the correctly rounded one and *first_failure the
 * lowest such bit pattern (0xffffffff if none).  FP_SWEEP_SQRT_STEP: the
 * step's square root of arguments in [2^-96, FLT_MAX], each result tested
 * exactly in binary64.  Diagnostic; no counterpart in the reference.
 * Blocking. */
#define FP_SWEEP_SQRT_STEP 0
int fp_math_sweep(fp_ctx *ctx, int fn, uint32_t first_bits, uint32_t last_bits, uint64_t *failures,
                  uint32_t *first_failure);

/* ---- timing ----------------------------------------------------------- */

/* HIP-event timing of each kernel launch on the context stream (the
 * reference's FLEXPART_GPU_PROFILE wall-clock sections, timeloop.rs:398-467,
 * with device timestamps).  kernel ids: FP_KERNEL_*. */
#define FP_KERNEL_STEP 0
#define FP_KERNEL_GRID 1
#define FP_KERNEL_COMPACT 2
#define FP_KERNEL_SORT 3
#define FP_KERNEL_PBL 4     /* fp_pbl_from_surface */
#define FP_KERNEL_OUTPUT 5  /* fp_output_sample */
#define FP_KERNEL_COUNT 6
int fp_profile_enable(fp_ctx *ctx, int enable);
/* Blocking: total milliseconds and launch count since the last reset. */
int fp_profile_read(fp_ctx *ctx, int kernel, double *total_ms, uint64_t *launches);
int fp_profile_reset(fp_ctx *ctx);

/* ---- tuning ----------------------------------------------------------- */

/* Launch shape of the step kernel (gpu/workgroup.rs runtime_workgroup_size):
 * block size (64..1024, multiple of 64) and particles per thread (1..16). */
int fp_set_launch(fp_ctx *ctx, uint32_t block_size, uint32_t particles_per_thread);

/* Autotuner (gpu/workgroup.rs:133-147,342-445: candidate sweep, warm-up runs,
 * median of measured runs, deterministic tie-break, per-device JSON cache),
 * extended to particles per thread and the Morton re-sort interval.  Trials
 * run fp_step (and fp_sort_morton) on a snapshot of the current particle
 * state and output grid, which are restored afterwards.  The chosen launch
 * shape is applied to the context; the re-sort interval is returned for the
 * caller's time loop (ForwardSpatialSortConfig::interval_steps). */
typedef struct fp_autotune_params {
    uint32_t warmup_steps;        /* per candidate (reference: 2) */
    uint32_t measured_steps;      /* per candidate, median taken (reference: 7) */
    uint32_t sort_bits;           /* bits_per_axis for re-sort trials */
    uint32_t n_sort_intervals;    /* candidates in sort_intervals (0 = keep the current) */
    uint32_t sort_intervals[8];   /* steps between re-sorts; 0 = never */
    uint32_t sort_trial_steps;    /* steps timed per re-sort candidate */
    const char *cache_path;       /* JSON cache file, NULL = no cache */
} fp_autotune_params;

typedef struct fp_autotune_result {
    uint32_t block_size;
    uint32_t particles_per_thread;
    uint32_t sort_interval;
    uint32_t from_cache;
    double step_ms;               /* median step time of the chosen launch shape */
    double step_ms_with_sort;     /* mean step time including amortised re-sorts */
} fp_autotune_result;

int fp_autotune(fp_ctx *ctx, const fp_step_params *base, const fp_autotune_params *params,
                fp_autotune_result *out);

/* ---- batched ensembles -------------------------------------------------
 * An ensemble runs many independent simulations (members) of unequal
 * particle counts on one vertical grid, each with its own particle range,
 * met bracket, Philox key and counter and output grid, and a shared PBL
 * state.  One launch per sync step advances every member (fp_ens_step),
 * optionally replayed from a captured hipGraph.  Each member's result is
 * exactly what fp_step gives for that member alone.  No counterpart in the
 * reference, which runs one simulation per driver. */
typedef struct fp_ens fp_ens;

typedef struct fp_ens_desc {
    uint32_t n_members;
    uint32_t met_nx, met_ny, met_nz;
    uint32_t n_species;
    uint32_t out_nx, out_ny, out_nz;  /* per-member output grid (0 = none) */
    const uint64_t *capacities;       /* n_members particle capacities */
} fp_ens_desc;

int fp_ens_create(int device_ordinal, const fp_ens_desc *desc, fp_ens **out);
int fp_ens_destroy(fp_ens *ens);
const char *fp_ens_last_error(const fp_ens *ens);
int fp_ens_sync(fp_ens *ens);
/* member slots [0, n) from AoS records; n <= that member's capacity */
int fp_ens_upload_member(fp_ens *ens, uint32_t member, const fp_particle *p, uint64_t n);
int fp_ens_download_member(fp_ens *ens, uint32_t member, fp_particle *out, uint64_t n);
/* member met bracket (layout of fp_met_upload_bracket), from host or device
 * pointers; the level heights (host) must be the same for every member */
int fp_ens_met_upload_member(fp_ens *ens, uint32_t member, const float *u0, const float *v0,
                             const float *w0, const float *u1, const float *v1, const float *w1,
                             const float *level_heights);
int fp_ens_met_upload_member_dev(fp_ens *ens, uint32_t member, const float *u0, const float *v0,
                                 const float *w0, const float *u1, const float *v1,
                                 const float *w1, const float *level_heights_host);
int fp_ens_pbl_upload(fp_ens *ens, const float *ustar, const float *wstar, const float *hmix,
                      const float *oli, const float *precip);
int fp_ens_grid_configure(fp_ens *ens, const fp_grid_params *params);
int fp_ens_grid_clear(fp_ens *ens);
int fp_ens_grid_download_member(fp_ens *ens, uint32_t member, uint32_t *counts, uint32_t *mass_fp);
/* Step parameters for the next fp_ens_step calls: base holds everything but
 * the per-member Philox state (keys: 2 per member, NULL = base->key;
 * counters: 4 per member at the first step) and the temporal weight, which
 * is (time_in_bracket + k * direction * dt) / bracket_seconds, clamped to
 * [0, 1], at clock step k (k = 0 at the first step after this call).
 * dispatch: per-member slot counts (NULL = capacities).  Rebuilds the
 * launch table; captured graphs are kept while the table's size and the
 * launch shape stay the same (they read the tables at run time). */
int fp_ens_configure(fp_ens *ens, const fp_step_params *base, const uint32_t *keys,
                     const uint32_t *counters, const uint64_t *dispatch, double time_in_bracket_s,
                     double bracket_s);
/* n_steps steps of every member; use_graph = 1 replays captured hipGraphs:
 * min(n_steps, 36) consecutive steps (an output interval, say) captured as
 * one graph that later calls of the same length replay again, the rest of
 * the call from a one-step graph.  A graph is captured again when the
 * launch table's size or the launch shape changed since its capture.
 * counters_out (4 per member, may be NULL) receives each member's counter
 * after the last step. */
int fp_ens_step(fp_ens *ens, uint32_t n_steps, int use_graph, uint32_t *counters_out);
/* Launch shape of the step; a configured ensemble keeps its configuration
 * and step clock (the launch tables are rebuilt, graphs captured again at
 * the next fp_ens_step). */
int fp_ens_set_launch(fp_ens *ens, uint32_t block_size, uint32_t particles_per_thread);
/* Morton re-sort of every member's dispatch range: each member ends in the
 * order fp_sort_morton gives for it alone (replaces ParticleStore::
 * sort_spatially, particles/mod.rs:649-699, per member).  The launch tables
 * and a captured graph stay valid. */
int fp_ens_sort_morton(fp_ens *ens, const fp_sort_params *params);
/* Ensemble autotuner (the reference's workgroup autotuner, gpu/workgroup.rs:
 * 342-445, over the ensemble's graph path): block size in {128, 256, 512} x
 * particles per thread in {1, 2, 4} by the median of measured_steps steps
 * after warmup_steps, then params->sort_intervals with the winning shape,
 * each over sort_trial_steps steps with the sorts amortised.  Requires a
 * configured ensemble; the particle state, grids and step clock are restored,
 * and the winning launch shape is set.  cache_path is not used. */
int fp_ens_autotune(fp_ens *ens, const fp_autotune_params *params, fp_autotune_result *out);
/* mean milliseconds per fp_ens step launch over the last fp_ens_step call
 * (HIP events on the ensemble stream). Blocking. */
int fp_ens_last_step_ms(fp_ens *ens, double *ms);

/* ---- differentiable runs (adjoint) ----------------------------------
 * A forward run of point sources through the timestep, and its hand-written
 * adjoint: the gradient of a receptor misfit
 *     L = sum_r (c_r - obs_r)^2,
 *     c_r = sum over observation times t and particles p of m_p(t) W(X_p(t), node_r)
 * (W: cloud-in-cell weights of the observation lattice: met grid columns in
 * x and y, obs_nz levels obs_dz_m apart in z) with respect to the emission
 * of every source (q_s, kg), its release position (x, y in grid units, z in
 * m) and the coefficients c_k of a wind perturbation
 *     u += sum_k c_k Bu_k(i, j),  v += sum_k c_k Bv_k(i, j)
 * added to every level and time level of the met.  Source s releases
 * particles_per_source[s] particles at its position at t = 0, each with mass
 * q_s / particles_per_source[s].  The backward pass replays the forward's
 * Philox draws (counter-based), restarts from a state checkpoint every
 * checkpoint_every steps, recomputes the steps in between (replaying the
 * saved survivor maps when particles are compacted), and propagates
 * the adjoint state back through deposition, the Langevin sub-steps with
 * reflection, the Hanna profiles, advection and the wind interpolation.
 * It is the pathwise (common random numbers) derivative.  No counterpart in
 * the reference. */
typedef struct fp_adj fp_adj;

typedef struct fp_adj_desc {
    uint32_t n_sources;
    uint32_t met_nx, met_ny, met_nz;
    uint32_t n_met_levels;            /* time levels, >= 2 */
    uint32_t n_modes;                 /* wind coefficients */
    uint32_t obs_nz;                  /* observation lattice levels, >= 2 */
    float obs_dz_m;                   /* spacing of the lattice levels */
    uint32_t n_steps;
    uint32_t obs_first_step;          /* observe after steps obs_first_step..n_steps (>= 1) */
    uint32_t checkpoint_every;        /* >= 1 */
    uint32_t flags;                   /* FP_ADJ_* */
    const uint32_t *particles_per_source; /* n_sources counts (>= 1 each) */
} fp_adj_desc;

/* compact the particles after every step (the reference's compaction: the
 * forward saves every step's survivor map and the backward carries the
 * adjoint back through it) */
#define FP_ADJ_COMPACT 1u
/* backward computes the emission gradient only (the adjoint of the masses:
 * deposition factors, survivor maps and observations; no position or wind
 * adjoint) */
#define FP_ADJ_EMISSION_ONLY 2u

int fp_adj_create(int device_ordinal, const fp_adj_desc *desc, fp_adj **out);
int fp_adj_destroy(fp_adj *adj);
const char *fp_adj_last_error(const fp_adj *adj);
/* met time levels (host, level-major: u[l*npts + p]; layout of
 * fp_met_upload_bracket per level) seconds_between_levels apart */
int fp_adj_met_upload(fp_adj *adj, const float *u, const float *v, const float *w,
                      const float *level_heights, double seconds_between_levels);
/* wind perturbation basis: n_modes fields of met_nx * met_ny (host, mode-major) */
int fp_adj_basis_upload(fp_adj *adj, const float *basis_u, const float *basis_v);
int fp_adj_pbl_upload(fp_adj *adj, const float *ustar, const float *wstar, const float *hmix,
                      const float *oli, const float *precip);
/* receptors as lattice nodes (i, j, l) (host, 3 per receptor) and their
 * observed values (may be NULL: zero) */
int fp_adj_receptors(fp_adj *adj, uint32_t n_receptors, const uint32_t *nodes, const double *observed);
/* Forward run.  base: physics of every step (dt, scales, key, counter at
 * step 0, sub-steps, deposition; alpha and grid_fused are ignored: the
 * temporal weight follows interpolation_alpha over the met levels).
 * emission: n_sources, position: 3 per source, coeff: n_modes.
 * receptor_values (n_receptors, may be NULL) and loss (may be NULL) on
 * return.  Blocking. */
int fp_adj_forward(fp_adj *adj, const fp_step_params *base, const float *emission,
                   const float *position, const float *coeff, double *receptor_values,
                   double *loss);
/* Gradient of the loss of the last forward run.  Blocking. */
int fp_adj_backward(fp_adj *adj, double *grad_emission, double *grad_position, double *grad_coeff);
/* particle state after the last forward run (the first n slots; after
 * compaction the active particles come first) */
int fp_adj_download_state(fp_adj *adj, fp_particle *out, uint64_t n);
/* active particles at the start of every step of the last forward run and
 * after its last step (n_steps + 1 values) */
int fp_adj_active_counts(fp_adj *adj, uint64_t *counts);
/* device milliseconds of the last forward and backward calls */
int fp_adj_last_ms(fp_adj *adj, double *forward_ms, double *backward_ms);

#ifdef __cplusplus
}
#endif

#endif /* FLEXPART_HIP_H */
