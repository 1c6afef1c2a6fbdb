// fp_static.hip - the entry point that allocates nothing: the survivor map of
// a compaction over caller-owned device buffers, with its tile counts in a
// library-owned static device array.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc) get the survivor map that
// fp_compact computes - ascending slot order of the active slots, and their
// count - without a context and without any allocation.  The only scratch the
// kernels need, one 32-bit count per 2048-slot tile, lives in `scratch`
// below; each device that runs the library holds its own copy until the
// process ends.  Every call on a device shares that device's copy, so calls
// that overlap on one device (two streams, two threads) race on it.
#include <mutex>

#include "fp_internal.hpp"

namespace {

constexpr uint32_t kScratchWords = 65536;  // 256 KiB: tile counts for 134,217,728 slots
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint32_t scratch[kScratchWords];

// Device address of `scratch` on each device, looked up once per device with
// that device current (host code never passes `scratch` itself to a kernel:
// on the host it names a stand-in, not the device array).
void *g_scratch_dev[kMaxDevices] = {};
std::once_flag g_scratch_once[kMaxDevices];
hipError_t g_scratch_status[kMaxDevices] = {};

hipError_t scratch_on(int device, uint32_t **out)
{
    std::call_once(g_scratch_once[device], [device] {
        int prev = 0;
        hipError_t e = hipGetDevice(&prev);
        if (e == hipSuccess) e = hipSetDevice(device);
        void *p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess && p == nullptr) e = hipErrorInvalidSymbol;
        const hipError_t r = hipSetDevice(prev);
        if (e == hipSuccess) e = r;
        g_scratch_status[device] = e;
        if (e == hipSuccess) g_scratch_dev[device] = p;
    });
    *out = static_cast<uint32_t *>(g_scratch_dev[device]);
    return g_scratch_status[device];
}

int fail(int code, const char *msg)
{
    fp::set_context_free_error(msg);
    return code;
}

}  // namespace

extern "C" int fp_survivor_map_static(int device_ordinal, const uint32_t *flags, uint64_t n,
                                      uint32_t *survivor_map, uint32_t *active_count, void *stream)
{
    if (device_ordinal < 0 || device_ordinal >= kMaxDevices) return fail(FP_E_INVALID, "device ordinal out of range");
    if (n > FP_SURVIVOR_MAP_STATIC_MAX)
        return fail(FP_E_INVALID, "fp_survivor_map_static: more slots than its static scratch covers "
                                  "(FP_SURVIVOR_MAP_STATIC_MAX = 134217728)");
    if (!active_count || (n && (!flags || !survivor_map))) return fail(FP_E_INVALID, "null argument");
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(FP_E_DEVICE, hipGetErrorString(e));
    if (dev != device_ordinal) return fail(FP_E_INVALID, "device_ordinal is not the current device");
    uint32_t *tiles = nullptr;
    e = scratch_on(device_ordinal, &tiles);
    if (e != hipSuccess) return fail(FP_E_DEVICE, hipGetErrorString(e));
    static_assert((uint64_t)kScratchWords * 2048u == FP_SURVIVOR_MAP_STATIC_MAX, "scratch covers the maximum");
    e = fp::launch_survivor_map_flags(flags, (uint32_t)n, tiles, survivor_map, active_count,
                                      static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(FP_E_DEVICE, hipGetErrorString(e));
    return FP_OK;
}
