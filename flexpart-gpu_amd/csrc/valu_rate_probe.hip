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
    uint32_t su = (uint32_t)smask;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        a[k] = 1e-3f * (float)(t + k) + 0.5f;
        d[k] = (double)a[k];
        q[k] = t * 7u + k;
        u[k] = t * 13u + k;
    }
    if (OP == 44) asm volatile("s_mov_b64 vcc, %0" : : "s"(smask) : "vcc");
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
            if (OP == 12) asm volatile("v_mov_b32_e32 %0, %1" : "=v"(a[k]) : "v"(a[(k + 1) & 7]));
            if (OP == 13) asm volatile("v_mov_b32_e32 %0, 0x3f7fbe77" : "=v"(a[k]));
            if (OP == 14) asm volatile("v_mov_b32_e32 %0, %1" : "=v"(a[k]) : "s"(sf));
            if (OP == 15) asm volatile("v_fmac_f32_e32 %0, %1, %2" : "+v"(a[k]) : "v"(vf), "v"(a[(k + 1) & 7]));
            if (OP == 16) asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(a[k]) : "v"(vf), "s"(sf));
            if (OP == 17) asm volatile("v_cndmask_b32_e32 %0, %0, %1, vcc" : "+v"(a[k]) : "v"(vf));
            if (OP == 18) asm volatile("v_sub_f32_e32 %0, %1, %0" : "+v"(a[k]) : "v"(vf));
            if (OP == 19) asm volatile("v_and_b32_e32 %0, %1, %0" : "+v"(u[k]) : "v"(u[(k + 1) & 7]));
            if (OP == 20) asm volatile("v_cmp_gt_f32_e32 vcc, %0, %1" : : "v"(a[k]), "v"(vf) : "vcc");
            if (OP == 21) asm volatile("v_med3_f32 %0, %0, %1, %2" : "+v"(a[k]) : "v"(vf), "v"(a[(k + 1) & 7]));
            if (OP == 22) asm volatile("v_mul_f64 %0, %0, %1" : "+v"(d[k]) : "v"(d[(k + 1) & 7]));
            if (OP == 23) asm volatile("v_cvt_f32_f64_e32 %0, %1" : "=v"(a[k]) : "v"(d[k]));
            if (OP == 24) asm volatile("v_exp_f32_e32 %0, %0" : "+v"(a[k]));
            if (OP == 25) {
                uint64_t c;
                asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(q[k]), "=s"(c) : "v"(u[k]), "v"(u[(k + 1) & 7]));
            }
            if (OP == 26) asm volatile("v_bitop3_b32 %0, %1, %2, %0 bitop3:0x96" : "+v"(u[k]) : "v"(u[(k + 1) & 7]), "v"(u[(k + 2) & 7]));
            if (OP == 27) asm volatile("v_fmamk_f32 %0, %0, 0x3e99999a, %1" : "+v"(a[k]) : "v"(vf));
            if (OP == 28) asm volatile("v_add_u32_e32 %0, %1, %0" : "+v"(u[k]) : "v"(u[(k + 1) & 7]));
            if (OP == 29) asm volatile("v_ldexp_f32 %0, %0, %1" : "+v"(a[k]) : "v"(u[k]));
            if (OP == 30) asm volatile("v_rndne_f32_e32 %0, %0" : "+v"(a[k]));
            if (OP == 31) asm volatile("v_sqrt_f32_e32 %0, %0" : "+v"(a[k]));
            if (OP == 32) {
                uint64_t m;
                asm volatile("v_cmp_class_f32_e64 %0, %1, %2" : "=s"(m) : "v"(a[k]), "v"(u[k]));
            }
            if (OP == 33) asm volatile("v_div_fixup_f32 %0, %0, %1, %2" : "+v"(a[k]) : "v"(vf), "v"(a[(k + 1) & 7]));
            if (OP == 34) asm volatile("v_mul_f32_e32 %0, %1, %0" : "+v"(a[k]) : "v"(vf));
            if (OP == 35) asm volatile("v_mov_b64_e32 %0, %1" : "=v"(d[k]) : "v"(d[(k + 1) & 7]));
            if (OP == 36) asm volatile("v_fma_f32 %0, %0, %1, 1.0" : "+v"(a[k]) : "v"(vf));
            if (OP == 37) asm volatile("v_cvt_i32_f32_e32 %0, %1" : "=v"(u[k]) : "v"(a[k]));
            if (OP == 38) asm volatile("v_cndmask_b32_e64 %0, %0, %1, vcc" : "+v"(a[k]) : "v"(vf));
            if (OP == 39) {
                uint64_t m;
                asm volatile("v_cmp_gt_f32_e64 %0, %1, %2" : "=s"(m) : "v"(a[k]), "v"(vf));
                asm volatile("s_nop 1\n\tv_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a[k]) : "v"(vf), "s"(m));
            }
            if (OP == 43) asm volatile("v_cmp_gt_f32_e32 vcc, %0, %1\n\ts_nop 1\n\tv_cndmask_b32_e32 %0, %0, %1, vcc"
                                       : "+v"(a[k]) : "v"(vf) : "vcc");
            if (OP == 44) asm volatile("v_cndmask_b32_e32 %0, %0, %1, vcc" : "+v"(a[k]) : "v"(vf));
            // a compare to VCC read by several selects (five instructions per 4 chain steps)
            if (OP == 45 && (k & 3) == 0)
                asm volatile("v_cmp_gt_f32_e32 vcc, %0, %4\n\ts_nop 1\n\tv_cndmask_b32_e32 %0, %0, %4, vcc\n\t"
                             "v_cndmask_b32_e32 %1, %1, %4, vcc\n\tv_cndmask_b32_e32 %2, %2, %4, vcc\n\t"
                             "v_cndmask_b32_e32 %3, %3, %4, vcc"
                             : "+v"(a[k]), "+v"(a[k + 1]), "+v"(a[k + 2]), "+v"(a[k + 3]) : "v"(vf) : "vcc");
            if (OP == 46 && (k & 3) == 0)
                asm volatile("v_cmp_gt_f32_e32 vcc, %0, %4\n\ts_nop 1\n\tv_cndmask_b32_e32 %0, %0, %4, vcc\n\t"
                             "v_add_f32_e32 %1, %4, %1\n\tv_add_f32_e32 %2, %4, %2\n\t"
                             "v_cndmask_b32_e32 %3, %3, %4, vcc"
                             : "+v"(a[k]), "+v"(a[k + 1]), "+v"(a[k + 2]), "+v"(a[k + 3]) : "v"(vf) : "vcc");
            if (OP == 47 && (k & 3) == 0)
                asm volatile("v_cmp_gt_f32_e64 %4, %0, %5\n\ts_nop 1\n\tv_cndmask_b32_e64 %0, %0, %5, %4\n\t"
                             "v_cndmask_b32_e64 %1, %1, %5, %4\n\tv_cndmask_b32_e64 %2, %2, %5, %4\n\t"
                             "v_cndmask_b32_e64 %3, %3, %5, %4"
                             : "+v"(a[k]), "+v"(a[k + 1]), "+v"(a[k + 2]), "+v"(a[k + 3]), "=&s"(q[k]) : "v"(vf));
            // one dependent chain per thread (no instruction-level parallelism)
            if (OP == 48) asm volatile("v_add_f32_e32 %0, %1, %0" : "+v"(a[0]) : "v"(vf));
            if (OP == 49) asm volatile("v_fmac_f32_e32 %0, %1, %0" : "+v"(a[0]) : "v"(vf));
            if (OP == 50) asm volatile("v_max_f32_e32 %0, %1, %0" : "+v"(a[0]) : "v"(vf));
            if (OP == 53) asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(a[0]) : "v"(vf), "v"(a[1]));
            if (OP == 54) asm volatile("v_mul_f32_e32 %0, %1, %0" : "+v"(a[0]) : "v"(vf));
            if (OP == 55) asm volatile("v_fmac_f32_e32 %0, %1, %0" : "+v"(a[k & 1]) : "v"(vf));
            if (OP == 56) asm volatile("v_fmac_f32_e32 %0, %1, %0" : "+v"(a[k & 3]) : "v"(vf));
            if (OP == 57) asm volatile("v_fmaak_f32 %0, %0, %1, 0x3e99999a" : "+v"(a[0]) : "v"(vf));
            if (OP == 58) asm volatile("v_fma_f64 %0, %0, %0, 0.5" : "+v"(d[0]));
            // two forms per chain step (counted as two instructions)
            if (OP == 51) asm volatile("v_add_f32_e32 %0, %2, %0\n\ts_add_u32 %1, %1, 0x3f"
                                       : "+v"(a[k]), "+s"(su) : "v"(vf) : "scc");
            if (OP == 52) asm volatile("v_add_f32_e32 %0, %1, %0\n\ts_nop 0" : "+v"(a[k]) : "v"(vf));
            if (OP == 40) asm volatile("v_add_f32_e32 %0, %2, %0\n\tv_max_f32_e32 %1, %2, %1"
                                       : "+v"(a[k]), "+v"(a[(k + 4) & 7]) : "v"(vf));
            if (OP == 41) asm volatile("v_add_f32_e32 %0, %2, %0\n\tv_mov_b32_e32 %1, %0"
                                       : "+v"(a[k]), "=v"(u[k]) : "v"(vf));
            if (OP == 42) asm volatile("v_add_f32_e32 %0, %2, %0\n\tv_mul_f32_e32 %1, %2, %1"
                                       : "+v"(a[k]), "+v"(a[(k + 4) & 7]) : "v"(vf));
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += a[k] + (float)d[k] + (float)q[k] + (float)u[k];
    s += (float)su;
    out[t] = s;
}

// per_step: instructions per chain step of the form (the 8 chains of a loop
// iteration issue 8 * per_step instructions)
template <int OP>
int run(const char *name, float *out, hipEvent_t e0, hipEvent_t e1, double per_step = 1.0)
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
    const double per_simd = (double)kBlocks * 4.0 * kIter * 8.0 * per_step / kSimds;
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
    rc |= run<12>("v_mov_b32_e32 v, v", out, e0, e1);
    rc |= run<13>("v_mov_b32_e32 v, literal", out, e0, e1);
    rc |= run<14>("v_mov_b32_e32 v, s", out, e0, e1);
    rc |= run<15>("v_fmac_f32_e32 v, v, v", out, e0, e1);
    rc |= run<16>("v_fma_f32 v, v, v, s", out, e0, e1);
    rc |= run<17>("v_cndmask_b32_e32 v, v, v, vcc", out, e0, e1);
    rc |= run<18>("v_sub_f32_e32 v, v, v", out, e0, e1);
    rc |= run<19>("v_and_b32_e32 v, v, v", out, e0, e1);
    rc |= run<20>("v_cmp_gt_f32_e32 vcc, v, v", out, e0, e1);
    rc |= run<21>("v_med3_f32 v, v, v, v", out, e0, e1);
    rc |= run<22>("v_mul_f64 v[2], v[2], v[2]", out, e0, e1);
    rc |= run<23>("v_cvt_f32_f64_e32 v, v[2]", out, e0, e1);
    rc |= run<24>("v_exp_f32_e32 v, v", out, e0, e1);
    rc |= run<25>("v_mad_u64_u32 v[2], s[2], v, v, v[2]", out, e0, e1);
    rc |= run<26>("v_bitop3_b32 v, v, v, v", out, e0, e1);
    rc |= run<27>("v_fmamk_f32 v, v, literal, v", out, e0, e1);
    rc |= run<28>("v_add_u32_e32 v, v, v", out, e0, e1);
    rc |= run<29>("v_ldexp_f32 v, v, v", out, e0, e1);
    rc |= run<30>("v_rndne_f32_e32 v, v", out, e0, e1);
    rc |= run<31>("v_sqrt_f32_e32 v, v", out, e0, e1);
    rc |= run<32>("v_cmp_class_f32_e64 s[2], v, v", out, e0, e1);
    rc |= run<33>("v_div_fixup_f32 v, v, v, v", out, e0, e1);
    rc |= run<34>("v_mul_f32_e32 v, v, v", out, e0, e1);
    rc |= run<35>("v_mov_b64_e32 v[2], v[2]", out, e0, e1);
    rc |= run<36>("v_fma_f32 v, v, v, inline constant", out, e0, e1);
    rc |= run<37>("v_cvt_i32_f32_e32 v, v", out, e0, e1);
    rc |= run<38>("v_cndmask_b32_e64 v, v, v, vcc", out, e0, e1);
    rc |= run<39>("v_cmp_gt_f32_e64 s[2] + v_cndmask_b32_e64 on it", out, e0, e1, 2.0);
    rc |= run<43>("v_cmp_gt_f32_e32 vcc + v_cndmask_b32_e32 on it", out, e0, e1, 2.0);
    rc |= run<44>("v_cndmask_b32_e32 v, v, v, vcc (vcc set)", out, e0, e1);
    rc |= run<45>("v_cmp_e32 vcc + 4 v_cndmask_b32_e32 on it", out, e0, e1, 1.25);
    rc |= run<46>("v_cmp_e32 vcc + cndmask_e32, 2 adds, cndmask_e32", out, e0, e1, 1.25);
    rc |= run<47>("v_cmp_e64 s[2] + 4 v_cndmask_b32_e64 on it", out, e0, e1, 1.25);
    rc |= run<48>("v_add_f32_e32, one dependent chain", out, e0, e1);
    rc |= run<49>("v_fmac_f32_e32, one dependent chain", out, e0, e1);
    rc |= run<50>("v_max_f32_e32, one dependent chain", out, e0, e1);
    rc |= run<53>("v_fma_f32 v, v, v, v, one dependent chain", out, e0, e1);
    rc |= run<54>("v_mul_f32_e32, one dependent chain", out, e0, e1);
    rc |= run<55>("v_fmac_f32_e32, two dependent chains", out, e0, e1);
    rc |= run<56>("v_fmac_f32_e32, four dependent chains", out, e0, e1);
    rc |= run<57>("v_fmaak_f32, one dependent chain", out, e0, e1);
    rc |= run<58>("v_fma_f64, one dependent chain", out, e0, e1);
    rc |= run<51>("v_add_f32 + s_add_u32 (pairs)", out, e0, e1, 2.0);
    rc |= run<52>("v_add_f32 + s_nop 0 (pairs)", out, e0, e1, 2.0);
    rc |= run<40>("v_add_f32 + v_max_f32 (pairs)", out, e0, e1, 2.0);
    rc |= run<41>("v_add_f32 + v_mov_b32 (pairs)", out, e0, e1, 2.0);
    rc |= run<42>("v_add_f32 + v_mul_f32 (pairs)", out, e0, e1, 2.0);
    CHK(hipFree(out));
    if (rc == 0) std::printf("probe done\n");
    return rc;
}
