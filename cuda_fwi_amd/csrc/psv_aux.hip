// psv_aux.hip -- the small kernels around the time step: FP64 misfit and
// residual, FP64 shot-sum of the gradients, fill and strided copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include "psv_aux.h"

namespace psv {

// Scratch of the no-allocation entry point: partial sums, per-shot misfits and
// the range flag of psvfwi_misfit_noalloc.  One copy per GPU, at its own address.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

hipError_t noalloc_scratch(void** ptr)
{
    constexpr int kMaxDevices = 64;
    static void* addr[kMaxDevices] = {};
    static std::mutex mu;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    return hipSuccess;
}

constexpr int kRedThreads = 256;

// One block per (shot slot, chunk).  Residual d - d_obs is formed in FP64
// (calculate_l2_adjoint_sources, CUDA_week2/fd_wave_model.cu:972-1002),
// written back in the plan's type as the shot's adjoint source, and
// 0.5 dt diff^2 is reduced in FP64 in a fixed order: each block writes one
// partial, misfit_finish adds the partials of a shot in chunk order.
template <typename T>
__global__ __launch_bounds__(kRedThreads) void misfit_partial_kernel(
    const T* syn, const double* obs, T* res, const size_t* off, const size_t* len,
    int nchunk, double half_dt, double* partial, int* flag)
{
    const int slot = blockIdx.x / nchunk, chunk = blockIdx.x % nchunk;
    const size_t n = len[slot], o = off[slot];
    const size_t per = (n + nchunk - 1) / nchunk;
    const size_t b = (size_t)chunk * per;
    const size_t e = b + per < n ? b + per : n;
    double acc = 0.0;
    for (size_t i = b + threadIdx.x; i < e; i += kRedThreads) {
        const double d = (double)syn[o + i] - obs[o + i];
        res[o + i] = (T)d;
        acc += half_dt * d * d;
        if (!(fabs(d) <= 1.0e10)) *flag = 1;  // plain store: every writer stores 1
    }
    __shared__ double red[kRedThreads];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void misfit_finish_kernel(const double* partial, int nslots, int nchunk, double* out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    double m = 0.0;
    for (int c = 0; c < nchunk; c++) m += partial[(size_t)s * nchunk + c];
    out[s] = m;
}

// sum over shots, in the order given by `order`, of the [3][nfz][nfx] gradients
template <typename T>
__global__ void grad_sum_kernel(const T* grad, size_t gst, const int* order, int nshots,
                                size_t n, double* out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int k = 0; k < nshots; k++) acc += (double)grad[(size_t)order[k] * gst + i];
    out[i] = acc;
}

template <typename T>
__global__ void fill_kernel(T* p, size_t n, T v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < n; k += step) p[k] = v;
}

template <typename T>
hipError_t launch_misfit(const T* syn, const double* obs, T* res, const size_t* off,
                         const size_t* len, int nslots, int nchunk, double half_dt,
                         double* partial, double* out, int* flag, hipStream_t s)
{
    hipLaunchKernelGGL((misfit_partial_kernel<T>), dim3(nslots * nchunk), dim3(kRedThreads), 0, s,
                       syn, obs, res, off, len, nchunk, half_dt, partial, flag);
    hipLaunchKernelGGL(misfit_finish_kernel, dim3((nslots + 63) / 64), dim3(64), 0, s, partial,
                       nslots, nchunk, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_grad_sum(const T* grad, size_t gst, const int* order, int nshots, size_t n,
                           double* out, hipStream_t s)
{
    hipLaunchKernelGGL((grad_sum_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       grad, gst, order, nshots, n, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fill(T* p, size_t n, T v, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((fill_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// blockIdx.y picks the copy; the blocks of x stride over it in 16-byte units
__global__ __launch_bounds__(256) void copies_kernel(const CopyList l)
{
    const int c = blockIdx.y;
    const uint4* src = (const uint4*)l.src[c];
    uint4* dst = (uint4*)l.dst[c];
    const size_t n = l.bytes[c] / 16;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}

hipError_t launch_copies(const CopyList& l, hipStream_t s)
{
    if (l.n <= 0) return hipSuccess;
    if (l.n > kMaxCopies) return hipErrorInvalidValue;
    size_t most = 0;
    for (int i = 0; i < l.n; i++) {
        if ((l.bytes[i] | (size_t)l.src[i] | (size_t)l.dst[i]) & 15) return hipErrorInvalidValue;
        if (l.bytes[i] > most) most = l.bytes[i];
    }
    size_t blocks = (most / 16 + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(copies_kernel, dim3((unsigned)blocks, (unsigned)l.n), dim3(256), 0, s, l);
    return hipGetLastError();
}

template hipError_t launch_misfit<float>(const float*, const double*, float*, const size_t*,
                                         const size_t*, int, int, double, double*, double*, int*,
                                         hipStream_t);
template hipError_t launch_misfit<double>(const double*, const double*, double*, const size_t*,
                                          const size_t*, int, int, double, double*, double*, int*,
                                          hipStream_t);
template hipError_t launch_grad_sum<float>(const float*, size_t, const int*, int, size_t, double*,
                                           hipStream_t);
template hipError_t launch_grad_sum<double>(const double*, size_t, const int*, int, size_t,
                                            double*, hipStream_t);
template hipError_t launch_fill<float>(float*, size_t, float, hipStream_t);
template hipError_t launch_fill<double>(double*, size_t, double, hipStream_t);
template hipError_t launch_fill<int>(int*, size_t, int, hipStream_t);
template hipError_t launch_misfit<dfloat>(const dfloat*, const double*, dfloat*, const size_t*,
                                          const size_t*, int, int, double, double*, double*, int*,
                                          hipStream_t);
template hipError_t launch_grad_sum<dfloat>(const dfloat*, size_t, const int*, int, size_t, double*,
                                            hipStream_t);
template hipError_t launch_fill<dfloat>(dfloat*, size_t, dfloat, hipStream_t);
template hipError_t launch_misfit<ffloat>(const ffloat*, const double*, ffloat*, const size_t*,
                                          const size_t*, int, int, double, double*, double*, int*,
                                          hipStream_t);
template hipError_t launch_grad_sum<ffloat>(const ffloat*, size_t, const int*, int, size_t, double*,
                                            hipStream_t);
template hipError_t launch_fill<ffloat>(ffloat*, size_t, ffloat, hipStream_t);

}  // namespace psv
