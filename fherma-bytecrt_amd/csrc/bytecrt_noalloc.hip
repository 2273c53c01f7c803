// Allocation-free product entry point: bcrt_run_device with the residue
// planes in a file-scope device array (`scratch`) instead of the engine's
// workspace, for callers that cannot allocate device memory (code under
// stream capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Each GPU holds its own copy of `scratch` from the first call on it until
// the process ends.  Host code never uses the array's name as an address: the
// device address of each GPU's copy is looked up once per device with
// hipGetSymbolAddress (that device current) and passed to the kernels.  The
// encode writes every residue word of the planes before any kernel reads one,
// so nothing in `scratch` has to start at zero; calls that overlap on one GPU
// share the array and race.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/bytecrt.h"

namespace bcrt {
// bytecrt_engine.cpp
int engine_device(const bcrt_engine* e);
int engine_fail(bcrt_engine* e, int code, const std::string& msg);
int run_with_planes(bcrt_engine* e, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, uint32_t* planes,
                    uint32_t batch, void* stream);
}  // namespace bcrt

namespace {

// 64 MiB: the residue planes (2 * P * N words) of one product at every
// geometry the plan accepts (P <= 128 primes, N <= 65,536), or of several
// smaller products (four at N = 32,768, W = 868).
constexpr size_t kScratchWords = size_t(16) << 20;
constexpr int kMaxDevices = 64;

std::mutex g_scratch_m;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {

// Device address of `device`'s copy of the scratch array, looked up on the
// first call for that device with the device current.
hipError_t scratch_on(int device, uint32_t** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_m);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" int bcrt_run_device_noalloc(bcrt_engine* e, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c,
                                       uint32_t batch, void* stream) {
    if (!e) return BCRT_EINVAL;
    if (!d_a || !d_b || !d_c) return bcrt::engine_fail(e, BCRT_EINVAL, "ByteCRT null device pointer");
    if (batch == 0) return bcrt::engine_fail(e, BCRT_EINVAL, "ByteCRT batch must be positive");
    const size_t words = bcrt_planes_words(e, batch);
    if (words > kScratchWords)
        return bcrt::engine_fail(e, BCRT_EINVAL,
                                 "ByteCRT no-alloc run: " + std::to_string(batch) + " products need " +
                                     std::to_string(words * 4) + " bytes of residue planes, above the scratch size of " +
                                     std::to_string(kScratchWords * 4));
    uint32_t* planes = nullptr;
    const hipError_t he = scratch_on(bcrt::engine_device(e), &planes);
    if (he != hipSuccess)
        return bcrt::engine_fail(e, BCRT_ERUNTIME, std::string("ByteCRT no-alloc run: scratch lookup: ") +
                                                       hipGetErrorString(he));
    return bcrt::run_with_planes(e, d_a, d_b, d_c, planes, batch, stream);
}
