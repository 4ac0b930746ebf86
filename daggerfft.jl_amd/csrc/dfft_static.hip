// dfft_static.hip -- dfft_execute_static: dfft_execute without any device
// allocation.  No counterpart in the reference; it is for callers that cannot
// allocate (HIP graph capture, tight serving loops).
//
// The only device memory a plan's execution needs beyond its input and output
// is the C2R workspace (the intermediate complex array of the passes before
// the last one).  Here that workspace is `scratch`, a file-scope device array
// of DFFT_STATIC_SCRATCH_BYTES.  Each device holds its own copy at its own
// address; host code never passes the array's name to a kernel (on the host
// it is only a stand-in for the symbol) but the device address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device on that device's first call.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/daggerfft_hip.h"
#include "dfft_plan.hpp"

namespace {

__device__ __align__(256) unsigned char scratch[DFFT_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices];  // scratch's device address per device; nullptr until looked up

// scratch's address on the current device `dev`.
int scratch_address(int dev, void **out) {
    if (dev < 0 || dev >= kMaxDevices) return DFFT_ERR_ARG;
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        if (hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)))
            return dfft::hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return dfft::hip_fail(hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = ptr;
    }
    *out = g_scratch[dev];
    return DFFT_OK;
}

}  // namespace

extern "C" int dfft_execute_static(const dfft_plan *p, const void *d_in, void *d_out, void *stream) {
    if (int rc = dfft::plan_check_io(p, d_in, d_out)) return rc;
    const size_t need = dfft::plan_work_bytes(p);
    if (need > DFFT_STATIC_SCRATCH_BYTES) {
        char msg[160];
        snprintf(msg, sizeof msg, "dfft_execute_static: workspace of %zu bytes exceeds the static scratch of %zu bytes",
                 need, (size_t)DFFT_STATIC_SCRATCH_BYTES);
        dfft::set_last_error(msg);
        return DFFT_ERR_SCRATCH;
    }
    int dev = -1;
    if (hipError_t e = hipGetDevice(&dev)) return dfft::hip_fail(e, "hipGetDevice");
    if (dev != dfft::plan_device(p)) return DFFT_ERR_ARG;
    void *work = nullptr;
    if (need)
        if (int rc = scratch_address(dev, &work)) return rc;
    return dfft::plan_enqueue(p, d_in, d_out, work, stream, false);
}
