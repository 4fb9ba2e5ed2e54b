// psv_aux.h -- launchers of the small kernels around the time step.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

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

}  // namespace psv
