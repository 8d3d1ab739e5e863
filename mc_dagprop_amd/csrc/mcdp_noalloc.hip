// mcdp_reverse_device_noalloc (C ABI in include/mcdp.h): the reverse walk of
// mcdp_reverse_device on a fixed scratch array instead of a workspace the
// library allocates, for callers that cannot allocate device memory (code
// under HIP graph capture, serving loops that forbid hipMalloc).
//
// The scratch is one file-scope device array per GPU, loaded with the code
// object and held until the process ends.  Its device address is taken once
// per device, with that device current, through hipGetSymbolAddress; host
// code uses only that address (the array's name is a host-side stand-in).
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (two streams, two threads) race on it.  Nothing in it needs to start at
// zero: every launch writes the partial rows, global slots, unit rows and
// group sums it reads.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/mcdp.h"
#include "mcdp_internal.hpp"

namespace {

constexpr size_t kScratchBytes = MCDP_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // per device; nullptr until looked up

// The scratch array's address on the current device (looked up once per
// device).
int scratch_address(void **out) {
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return mcdp::hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[(size_t)device]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return mcdp::hip_fail(e, "hipGetSymbolAddress (reverse-walk scratch)");
        if (!ptr) return mcdp::fail(MCDP_E_DEVICE, "hipGetSymbolAddress returned no address for the scratch");
        g_scratch_ptr[(size_t)device] = ptr;
    }
    *out = g_scratch_ptr[(size_t)device];
    return MCDP_OK;
}

}  // namespace

extern "C" {

int mcdp_reverse_device_noalloc(mcdp_propagator *prop, const int32_t *seeds_dev, const double *weights_dev,
                                const double *realized_dev, const double *durations_dev, const int32_t *cause_dev,
                                int64_t n, int64_t ld, int accumulate, double *objective_dev, double *grad_dmin_dev,
                                double *grad_earliest_dev, void *stream) {
    if (!prop) return mcdp::fail(MCDP_E_INVALID, "propagator is null");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return mcdp::fail(MCDP_E_DEVICE, "no HIP device available: gradients run on the GPU only");
    void *work = nullptr;
    int rc = scratch_address(&work);
    if (rc) return rc;
    return mcdp::reverse_run(prop, seeds_dev, weights_dev, realized_dev, durations_dev, cause_dev, n, ld, accumulate,
                             objective_dev, grad_dmin_dev, grad_earliest_dev, work, kScratchBytes,
                             static_cast<hipStream_t>(stream));
}

}  // extern "C"
