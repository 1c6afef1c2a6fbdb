// fp_surface.hip - PBL state from the surface fields, on the device.
//
// The step before the particle step (timeloop.rs:961-993): the surface
// fields of the two met levels of the bracket are interpolated to the current
// time (interpolate_surface_fields_linear, io/temporal.rs:122-214) and the
// PBL diagnostics of shaders/pbl_diagnostics.wgsl:220-305 derive u*, w*,
// hmix and 1/L from them.  The reference interpolates on the CPU and uploads
// the interpolated fields every step; here both bracket ends stay resident
// in HBM and one kernel per cell does the interpolation, the diagnostics and
// the write of the step kernel's two PBL records ({u*, w*, hmix, 1/L} and
// pbl_cell_terms_dev), so a step reads 80 B and writes 32 B per cell and
// nothing crosses PCIe.
//
// Transcendentals (log of the log-law, the cube root of w*) come from
// fp_math.h and the operation order is the shader's, so the values are those
// of the CPU oracle bit for bit.
#include "fp_physics.hpp"

namespace fp {
namespace {

constexpr uint32_t kSurfBlock = 256;

// constants module (lib.rs:24-28) and pbl_diagnostics.wgsl:20-26
constexpr float kRAir = 287.05f;
constexpr float kGa = 9.81f;
constexpr float kCpa = 1004.6f;
constexpr float kKarman = 0.4f;
constexpr float kNeutralOli = 1.0e-5f;

struct SurfArgs {
    const float *s0[FP_SURF_COUNT];
    const float *s1[FP_SURF_COUNT];
    float alpha;
    uint32_t n;
    fp_pbl_options o;
    float4 *pbl;
    float *pass;  // 3 * n or null
};

__device__ __forceinline__ bool sfinite(float v) { return v == v && fabsf(v) <= F32_MAX; }
__device__ __forceinline__ float spositive(float v, float fb) { return sfinite(v) && v > 0.0f ? v : fb; }
__device__ __forceinline__ float snonneg(float v, float fb) { return sfinite(v) && v >= 0.0f ? v : fb; }
__device__ __forceinline__ float sfin(float v, float fb) { return sfinite(v) ? v : fb; }

// estimate_friction_velocity (pbl_diagnostics.wgsl:130-152)
__device__ __forceinline__ float friction_velocity(float stress, float rho_in, float u10, float v10, float z0_in,
                                                   float zref_in)
{
    const float rho = spositive(rho_in, 1.225f);
    const float st = snonneg(stress, 0.0f);
    if (st > 0.0f) return fmaxf(fpm_sqrtf(st / rho), USTAR_MIN);
    const float z0 = spositive(z0_in, 0.1f);
    const float zr = fmaxf(spositive(zref_in, 10.0f), 1.01f * z0);
    const float ws = fmaxf(fpm_sqrtf(u10 * u10 + v10 * v10), 0.0f);
    if (ws <= 0.0f) return USTAR_MIN;
    const float us = kKarman * ws / fmaxf(fpm_logf(zr / z0), 1.0e-6f);
    return fmaxf(spositive(us, USTAR_MIN), USTAR_MIN);
}

// obukhov_length_from_flux (pbl_diagnostics.wgsl:160-187)
__device__ __forceinline__ float obukhov_from_flux(float ust_in, float t_in, float rho_in, float hf_in, float thr)
{
    const float h = sfin(hf_in, 0.0f);
    if (fabsf(h) < fmaxf(thr, 0.0f)) return __builtin_inff();
    const float u = fmaxf(spositive(ust_in, USTAR_MIN), USTAR_MIN);
    const float t = spositive(t_in, 300.0f);
    const float d = spositive(rho_in, 1.225f);
    const float num = -(d * kCpa * t * u * u * u);
    const float den = kKarman * kGa * h;
    if (fabsf(den) < 1.0e-9f) return __builtin_inff();
    const float ol = num / den;
    return (!sfinite(ol) || fabsf(ol) > 1.0e9f) ? __builtin_inff() : ol;
}

// convective_velocity_scale (pbl_diagnostics.wgsl:204-218)
__device__ __forceinline__ float convective_velocity(float hf, float rho_in, float t_in, float hmix_in, float hmin)
{
    if (hf <= 0.0f) return 0.0f;
    const float d = spositive(rho_in, 1.225f);
    const float t = spositive(t_in, 300.0f);
    const float h = spositive(hmix_in, hmin);
    const float bf = hf / (d * kCpa) * kGa / t;
    if (bf <= 0.0f || !sfinite(bf)) return 0.0f;
    return fmaxf(fpm_powf(bf * h, 1.0f / 3.0f), 0.0f);
}

__global__ void __launch_bounds__(kSurfBlock) pbl_surface_kernel(SurfArgs a)
{
    const uint32_t c = blockIdx.x * kSurfBlock + threadIdx.x;
    if (c >= a.n) return;
    // lerp of io/temporal.rs:300-303, with the exact end copies of :134-139
    float f[FP_SURF_COUNT];
    const float al = a.alpha, ua = 1.0f - al;
#pragma unroll
    for (int k = 0; k < FP_SURF_COUNT; ++k) {
        const float x0 = a.s0[k][c], x1 = a.s1[k][c];
        f[k] = al <= 0.0f ? x0 : (al >= 1.0f ? x1 : fmaf(x0, ua, x1 * al));
    }
    const fp_pbl_options &o = a.o;
    const float temp = spositive(f[FP_SURF_T2M], 300.0f);
    const float pres = spositive(f[FP_SURF_PRESSURE], 101325.0f);
    const float rho = pres / (kRAir * temp);
    const float est = friction_velocity(f[FP_SURF_STRESS], rho, f[FP_SURF_U10], f[FP_SURF_V10], o.roughness_length_m,
                                        o.wind_reference_height_m);
    const float us_in = f[FP_SURF_USTAR];
    const float ust = sfinite(us_in) && us_in > 0.0f ? fmaxf(us_in, USTAR_MIN) : fmaxf(est, USTAR_MIN);
    const float ol = obukhov_from_flux(ust, temp, rho, f[FP_SURF_SSHF], o.heat_flux_neutral_threshold_w_m2);
    const float oli_in = f[FP_SURF_OLI];
    float oli;
    if (sfinite(oli_in) && fabsf(oli_in) >= kNeutralOli) oli = oli_in;
    else oli = (!sfinite(ol) || fabsf(ol) > 1.0e5f) ? 0.0f : 1.0f / ol;  // inverse_obukhov_from_length
    // clamp_mixing_height: max(clamp(v, lo, hi), lo)
    const float h_in = f[FP_SURF_HMIX];
    const float hsel = sfinite(h_in) && h_in > 0.0f ? h_in : o.fallback_mixing_height_m;
    const float hmix = fmaxf(fminf(fmaxf(hsel, o.hmix_min_m), o.hmix_max_m), o.hmix_min_m);
    const float sshf = sfin(f[FP_SURF_SSHF], 0.0f);
    const float wst = convective_velocity(sshf, rho, temp, hmix, o.hmix_min_m);
    const float4 rec = make_float4(ust, wst, hmix, oli);
    a.pbl[2u * c] = rec;
    a.pbl[2u * c + 1u] = pbl_cell_terms_dev(rec);
    if (a.pass) {
        a.pass[c] = sshf;
        a.pass[a.n + c] = snonneg(f[FP_SURF_SSR], 0.0f);
        a.pass[2u * a.n + c] = snonneg(f[FP_SURF_STRESS], 0.0f);
    }
}

}  // namespace

hipError_t launch_pbl_surface(const float *const s0[FP_SURF_COUNT], const float *const s1[FP_SURF_COUNT], float alpha,
                              const fp_pbl_options &o, float4 *pbl, float *pass, uint32_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    SurfArgs a;
    for (int k = 0; k < FP_SURF_COUNT; ++k) {
        a.s0[k] = s0[k];
        a.s1[k] = s1[k];
    }
    a.alpha = alpha;
    a.n = n;
    a.o = o;
    a.pbl = pbl;
    a.pass = pass;
    pbl_surface_kernel<<<(n + kSurfBlock - 1) / kSurfBlock, kSurfBlock, 0, st>>>(a);
    return hipGetLastError();
}

}  // namespace fp
