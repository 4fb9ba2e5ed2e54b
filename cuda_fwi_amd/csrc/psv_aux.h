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

}  // namespace psv
