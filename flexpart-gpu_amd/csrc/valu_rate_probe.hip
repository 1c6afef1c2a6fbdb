// valu_rate_probe.hip - issue cost of single VALU instruction forms on the
// local GPU, the measured basis of the step kernel's instruction-issue bound.
//
// Each form runs as 8 independent dependency chains per thread, 1024 x 8
// instructions per thread, 20 wavefronts per SIMD over the launch (5120
// blocks of 256 threads on 1024 SIMDs).  The instruction is written in inline
// assembly so that its encoding and operand kinds are exactly the ones named
// (arithmetic only: no memory access in assembly).  Reported: nanoseconds per
// wavefront instruction per SIMD, best of two timed launches after a warm-up
// launch.  Build: flexpart-gpu_amd/build.py; run: flexpart-gpu_amd/lib/valu_rate_probe.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CHK(x)                                                                              \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);      \
            return 1;                                                                       \
        }                                                                                   \
    } while (0)

namespace {

constexpr int kIter = 1024;
constexpr int kBlocks = 5120;
constexpr int kSimds = 1024;

template <int OP>
__global__ __launch_bounds__(256) void probe(float *out, float sf, uint64_t smask)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    float a[8];
    double d[8];
    uint64_t q[8];
    uint32_t u[8];
    const float vf = sf + (float)(t & 1u) * 1e-30f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        a[k] = 1e-3f * (float)(t + k) + 0.5f;
        d[k] = (double)a[k];
        q[k] = t * 7u + k;
        u[k] = t * 13u + k;
    }
    for (int i = 0; i < kIter; ++i) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (OP == 0) asm volatile("v_add_f32_e32 %0, %1, %0" : "+v"(a[k]) : "v"(vf));
            if (OP == 1) asm volatile("v_mul_f32_e32 %0, 0x3f7fbe77, %0" : "+v"(a[k]));
            if (OP == 2) asm volatile("v_fmaak_f32 %0, %0, %1, 0x3e99999a" : "+v"(a[k]) : "v"(vf));
            if (OP == 3) asm volatile("v_add_f32_e32 %0, %1, %0" : "+v"(a[k]) : "s"(sf));
            if (OP == 4) asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(a[k]) : "v"(vf), "v"(a[(k + 1) & 7]));
            if (OP == 5) asm volatile("v_max_f32_e32 %0, %1, %0" : "+v"(a[k]) : "v"(vf));
            if (OP == 6) asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a[k]) : "v"(vf), "s"(smask));
            if (OP == 7) asm volatile("v_lshlrev_b32_e32 %0, 1, %0" : "+v"(u[k]));
            if (OP == 8) asm volatile("v_cvt_f32_u32_e32 %0, %1" : "=v"(a[k]) : "v"(u[k]));
            if (OP == 9) asm volatile("v_fma_f64 %0, %0, %0, 0.5" : "+v"(d[k]));
            if (OP == 10) asm volatile("v_lshl_add_u64 %0, %0, 2, %1" : "+v"(q[k]) : "v"(q[(k + 1) & 7]));
            if (OP == 11) asm volatile("v_rcp_f32_e32 %0, %0" : "+v"(a[k]));
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += a[k] + (float)d[k] + (float)q[k] + (float)u[k];
    out[t] = s;
}

template <int OP>
int run(const char *name, float *out, hipEvent_t e0, hipEvent_t e1)
{
    std::printf("# %s\n", name);
    std::fflush(stdout);
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        CHK(hipEventRecord(e0));
        hipLaunchKernelGGL(probe<OP>, dim3(kBlocks), dim3(256), 0, 0, out, 1.0e-3f, 0x5555555555555555ull);
        CHK(hipGetLastError());
        CHK(hipEventRecord(e1));
        CHK(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHK(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0 && ms < best) best = ms;
    }
    const double per_simd = (double)kBlocks * 4.0 * kIter * 8.0 / kSimds;
    std::printf("%-34s %8.3f ms  %6.3f ns per wavefront instruction per SIMD\n", name, best,
                best * 1e6 / per_simd);
    return 0;
}

}  // namespace

int main()
{
    float *out = nullptr;
    CHK(hipMalloc(&out, (size_t)kBlocks * 256 * sizeof(float)));
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    int rc = 0;
    rc |= run<0>("v_add_f32_e32 v, v, v", out, e0, e1);
    rc |= run<1>("v_mul_f32_e32 v, literal, v", out, e0, e1);
    rc |= run<2>("v_fmaak_f32 v, v, v, literal", out, e0, e1);
    rc |= run<3>("v_add_f32_e32 v, s, v", out, e0, e1);
    rc |= run<4>("v_fma_f32 v, v, v, v", out, e0, e1);
    rc |= run<5>("v_max_f32_e32 v, v, v", out, e0, e1);
    rc |= run<6>("v_cndmask_b32_e64 v, v, v, s[pair]", out, e0, e1);
    rc |= run<7>("v_lshlrev_b32_e32 v, 1, v", out, e0, e1);
    rc |= run<8>("v_cvt_f32_u32_e32 v, v", out, e0, e1);
    rc |= run<9>("v_fma_f64 v[2], v[2], v[2], 0.5", out, e0, e1);
    rc |= run<10>("v_lshl_add_u64 v[2], v[2], 2, v[2]", out, e0, e1);
    rc |= run<11>("v_rcp_f32_e32 v, v", out, e0, e1);
    CHK(hipFree(out));
    if (rc == 0) std::printf("probe done\n");
    return rc;
}
