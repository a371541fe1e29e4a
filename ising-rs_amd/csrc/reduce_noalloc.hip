// reduce_ising without allocation: the partial sums live in a file-scope
// device array instead of memory allocated per call, for callers that may not
// call hipMalloc (code under HIP graph capture, tight loops, harnesses that
// forbid it).  Same kernels as ising_reduce_ising (reduce.hip).
//
// The array is one per device and is shared by every call on that device:
// calls that overlap on one device race on it, so a caller serialises them
// (one stream, or waits between streams).  Each call writes every partial it
// reads, so nothing in it needs clearing between calls.
#include <mutex>

#include "../../include/ising_hip.h"
#include "kernels.h"

namespace ising {

// Largest lattice: n^3 <= kNoAllocBlocks * kReduceTile = 2^28 sites (n <= 645).
constexpr int kNoAllocBlocks = 65536;
// [0, 2): the totals (bond sum, spin sum); [2, 2 + 2 * kNoAllocBlocks): block partials.  1 MiB.
static __device__ __align__(256) int64_t scratch[2 + 2 * kNoAllocBlocks];

// E total = -j B and M total as doubles, from the int64 totals.
__global__ void noalloc_totals_kernel(const int64_t* bm, float j, double* out) {
    if (threadIdx.x == 0) {
        out[0] = -(double)j * (double)bm[0];
        out[1] = (double)bm[1];
    }
}

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
int64_t* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device (null until looked up)

// Device address of this device's copy of `scratch`, looked up once per device
// with that device current.
int scratch_of_current_device(int64_t** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(ISING_E_HIP, hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return set_error(ISING_E_HIP, "device index beyond the supported range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(ISING_E_HIP, e != hipSuccess ? hipGetErrorString(e) : "scratch symbol has no address");
        g_scratch[dev] = (int64_t*)p;
    }
    *out = g_scratch[dev];
    return ISING_OK;
}

}  // namespace
}  // namespace ising

using namespace ising;

extern "C" int ising_reduce_ising_noalloc(const int8_t* d_spins, int32_t n, float j, double* d_out, void* stream) {
    if (!d_spins || !d_out) return set_error(ISING_E_ARG, "null argument");
    if (n < 1 || (int64_t)n * n * n > (int64_t)kNoAllocBlocks * kReduceTile)
        return set_error(ISING_E_ARG, "n must be in 1..645 (n^3 <= 2^28 sites: the fixed partial-sum buffer)");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return set_error(ISING_E_HIP, "no HIP device visible");
    int64_t* sc = nullptr;
    const int rc = scratch_of_current_device(&sc);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = launch_reduce_ising(d_spins, n, sc + 2, sc, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(noalloc_totals_kernel, dim3(1), dim3(64), 0, st, (const int64_t*)sc, j, d_out);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return set_error(ISING_E_HIP, hipGetErrorString(e));
    return ISING_OK;
}
