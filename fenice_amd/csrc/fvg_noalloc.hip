// Entry points that allocate nothing (include/fvg.h: fvg_conserved_totals_noalloc). Their scratch memory is a
// file-scope device array; every GPU holds its own copy at its own address, looked up once per device with
// hipGetSymbolAddress. The array's name in host code is only a stand-in for the symbol, never a device address:
// kernels and copies get the looked-up pointer.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/fvg.h"
#include "fvg_host.hpp"

namespace fvg {
template <int NV, int WHAT> __global__ void k_totals_partial(const double*, long long, double*);
__global__ void k_totals_final(const double*, int, long long, int, double*);
}   // namespace fvg

using namespace fvg;

// the per-block partial sums of the conserved totals: 3 doubles per block of the first pass
static_assert(FVG_NOALLOC_TOTALS_SCRATCH == 3 * kTotalsBlocks * sizeof(double), "scratch size of the totals");
static __device__ __align__(256) double scratch[3 * kTotalsBlocks];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_lock;
double* g_scratch[kMaxDevices] = {};   // the device address of `scratch` on each device, once looked up

// this device's address of the scratch array (looked up on first use with the device current; the caller's current
// device is restored)
int scratch_of(int device, double** out) {
    std::lock_guard<std::mutex> guard(g_lock);
    if (g_scratch[device]) {
        *out = g_scratch[device];
        return FVG_OK;
    }
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) return FVG_EHIP;
    if (hipSetDevice(device) != hipSuccess) return FVG_EHIP;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    const hipError_t r = hipSetDevice(prev);
    if (e != hipSuccess || r != hipSuccess || !ptr) return FVG_EHIP;
    g_scratch[device] = (double*)ptr;
    *out = g_scratch[device];
    return FVG_OK;
}

}   // namespace

extern "C" int fvg_conserved_totals_noalloc(int device, const double* U, long long ncells, int ncomp, double* out, void* stream) {
    if (!U || !out) return FVG_EINVAL;
    if (ncomp != 4 && ncomp != 5) return FVG_EINVAL;
    if (ncells < 1 || ncells > FVG_NOALLOC_MAX_CELLS) return FVG_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) return FVG_EHIP;
    if (device < 0 || device >= ndev || device >= kMaxDevices) return FVG_EINVAL;
    double* part = nullptr;
    int rc = scratch_of(device, &part);
    if (rc) return rc;
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) return FVG_EHIP;
    if (prev != device && hipSetDevice(device) != hipSuccess) return FVG_EHIP;
    hipStream_t s = (hipStream_t)stream;
    // every block writes its three partial sums, so nothing has to be cleared between calls
    if (ncomp == 5) hipLaunchKernelGGL((k_totals_partial<5, 0>), dim3(kTotalsBlocks), dim3(256), 0, s, U, ncells, part);
    else hipLaunchKernelGGL((k_totals_partial<4, 0>), dim3(kTotalsBlocks), dim3(256), 0, s, U, ncells, part);
    hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(1), 0, s, part, kTotalsBlocks, ncells, 0, out);
    rc = hipGetLastError() == hipSuccess ? FVG_OK : FVG_EHIP;
    if (prev != device && hipSetDevice(prev) != hipSuccess) rc = FVG_EHIP;
    return rc;
}
