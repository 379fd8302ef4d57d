// culof_lof_device_noalloc: the device-resident forward for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc). No reference counterpart.
//
// Every scratch buffer of the call is carved from one fixed array, `scratch`
// below: CULOF_NOALLOC_SCRATCH_BYTES (256 MiB) per device, held from the
// library's first use of that device until the process ends. The pipeline is
// culof_lof_device's with its low-memory choices: the list sweep in place of
// the collect pipeline (whose buffers grow with the rows), query rows tiled to
// what the array holds, and the FP64 tie-break's brute-force scan. Results are
// bit-identical to culof_lof_device's. Every call on a device shares that
// device's array and leaves its contents behind (each call clears the counters
// it needs), so calls that overlap in time on one device race.

#include "internal.hpp"

#include "culof_c.h"

#include <mutex>

namespace culof {
namespace detail {

namespace {

constexpr size_t kScratchBytes = CULOF_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "whole 256-byte blocks");
constexpr int kMaxDevices = 64;

// One copy per device. In host code the name stands for a host-side shadow,
// not a device address: host code reaches the array only through the
// address scratch_region() looks up.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_region_mu;
void* g_region[kMaxDevices] = {};

// Device address of the current device's `scratch`, looked up once per
// device with that device current.
void* scratch_region() {
    int device = 0;
    CULOF_CHECK(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) throw CudaError("device index beyond the no-allocation table");
    std::lock_guard<std::mutex> lock(g_region_mu);
    if (!g_region[device]) {
        void* ptr = nullptr;
        CULOF_CHECK(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw CudaError("hipGetSymbolAddress returned no address for the scratch array");
        g_region[device] = ptr;
    }
    return g_region[device];
}

}  // namespace
}  // namespace detail
}  // namespace culof

extern "C" int culof_lof_device_noalloc(const float* points, int n_points, int n_dims, int k,
                                        int normalize, int form, float* scores, int* nbr_idx,
                                        float* nbr_d2, float* kdist2, float* lrd, void* stream) {
    using namespace culof::detail;
    return run_guarded([&] {
        void* region = scratch_region();
        forward_in_region(points, n_points, n_dims, k, normalize != 0, form, scores, nbr_idx, nbr_d2,
                          kdist2, lrd, region, kScratchBytes, static_cast<hipStream_t>(stream));
    });
}
