// psv_aux.h -- launchers of the small kernels around the time step.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "psv_dfloat.h"

namespace psv {

template <typename T>
hipError_t launch_misfit(const T* syn, const double* obs, T* res, const size_t* off,
                         const size_t* len, int nslots, int nchunk, double half_dt,
                         double* partial, double* out, int* flag, hipStream_t s);

template <typename T>
hipError_t launch_grad_sum(const T* grad, size_t gst, const int* order, int nshots, size_t n,
                           double* out, hipStream_t s);

template <typename T>
hipError_t launch_fill(T* p, size_t n, T v, hipStream_t s);

// Up to kMaxCopies device-to-device copies in one launch (a state save or
// restore: 5 wavefields and 8 CPML arrays).  Every size and address must be
// a multiple of 16 bytes.
constexpr int kMaxCopies = 16;
struct CopyList {
    const void* src[kMaxCopies];
    void* dst[kMaxCopies];
    size_t bytes[kMaxCopies];
    int n;
};
hipError_t launch_copies(const CopyList& l, hipStream_t s);

// Device scratch of the no-allocation misfit (psvfwi_misfit_noalloc): a static
// device array of kNoAllocBytes per GPU, held until the process ends.  Returns
// its address on the current device (resolved once per device with
// hipGetSymbolAddress; the array's host-side name is not a device address).
// Every call on one GPU shares it: calls that overlap on one GPU race.
constexpr size_t kNoAllocBytes = 544768;  // 1024 shots x (64 partials + 1 misfit) doubles + flag
hipError_t noalloc_scratch(void** ptr);

}  // namespace psv
