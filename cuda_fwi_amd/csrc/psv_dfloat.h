// psv_dfloat.h -- compensated FP32 arithmetic: a value held as the unevaluated
// sum hi + lo of two floats (a "double-word" or double-float number, Dekker
// 1971), with about 48 significant bits, computed with FP32 operations only.
//
// The step kernels are templates over their arithmetic type; instantiated with
// dfloat they run the same operation sequence with every +, -, *, / replaced by
// the double-word algorithms of Joldes, Muller and Popescu, "Tight and rigorous
// error bounds for basic building blocks of double-word arithmetic", ACM TOMS
// 44(2), 2017: AccurateDWPlusDW (Alg. 6, relative error below 3u^2 + 13u^3),
// DWTimesDW with FMA (Alg. 12, below 5u^2), DWDivDW (Alg. 17, below 15u^2 +
// 56u^3), u = 2^-24.  The error-free transforms they build on are TwoSum
// (Knuth), Fast2Sum (Dekker) and the FMA-based TwoProd.  They rely on every
// FP32 operation being rounded on its own: the library is compiled with
// contraction off, and the one fused multiply-add each product needs is
// written explicitly.
//
// Use: a plan of precision PSVFWI_FP32X2 (state, model and traces in dfloat,
// 8 bytes per value).  It resolves small differences of large wavefields (a
// residual 1e-3 .. 1e-4 of the direct wave) that plain FP32 rounds away.
#pragma once

#include <hip/hip_runtime.h>

namespace psv {

struct __align__(8) dfloat {
    float hi, lo;

    dfloat() = default;
    __host__ __device__ constexpr dfloat(float h, float l) : hi(h), lo(l) {}
    __host__ __device__ dfloat(double d) : hi((float)d), lo((float)(d - (double)(float)d)) {}
    __host__ __device__ dfloat(float f) : hi(f), lo(0.0f) {}
    __host__ __device__ dfloat(int i) : dfloat((double)i) {}
    __host__ __device__ explicit operator double() const { return (double)hi + (double)lo; }
};

namespace df {

__host__ __device__ inline float fma_(float a, float b, float c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fmaf(a, b, c);
#else
    return __builtin_fmaf(a, b, c);
#endif
}

// s + e = a + b exactly (Knuth's TwoSum, no condition on a, b)
__host__ __device__ inline void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// s + e = a + b exactly when |a| >= |b| or a = 0 (Dekker's Fast2Sum)
__host__ __device__ inline void fast_two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    e = b - (s - a);
}

}  // namespace df

// AccurateDWPlusDW
__host__ __device__ inline dfloat operator+(dfloat x, dfloat y)
{
    float sh, sl, th, tl;
    df::two_sum(x.hi, y.hi, sh, sl);
    df::two_sum(x.lo, y.lo, th, tl);
    sl = sl + th;
    df::fast_two_sum(sh, sl, sh, sl);
    sl = sl + tl;
    df::fast_two_sum(sh, sl, sh, sl);
    return dfloat(sh, sl);
}

// a normalized pair (|lo| <= ulp(hi) / 2, as every operation below returns it) is equal
// to another exactly when both words are
__host__ __device__ inline bool operator==(dfloat x, dfloat y) { return x.hi == y.hi && x.lo == y.lo; }
__host__ __device__ inline bool operator!=(dfloat x, dfloat y) { return !(x == y); }

__host__ __device__ inline dfloat operator-(dfloat x) { return dfloat(-x.hi, -x.lo); }
__host__ __device__ inline dfloat operator-(dfloat x, dfloat y) { return x + (-y); }

// DWTimesDW (with FMA)
__host__ __device__ inline dfloat operator*(dfloat x, dfloat y)
{
    const float ch = x.hi * y.hi;
    float cl1 = df::fma_(x.hi, y.hi, -ch);
    const float tl0 = x.lo * y.lo;
    const float tl1 = df::fma_(x.hi, y.lo, tl0);
    const float cl2 = df::fma_(x.lo, y.hi, tl1);
    cl1 = cl1 + cl2;
    float zh, zl;
    df::fast_two_sum(ch, cl1, zh, zl);
    return dfloat(zh, zl);
}

// DWTimesFP (x * y, y a float), used by the division
__host__ __device__ inline dfloat times_fp(dfloat x, float y)
{
    const float ch = x.hi * y;
    const float cl1 = df::fma_(x.hi, y, -ch);
    const float cl2 = x.lo * y;
    float th, tl1;
    df::fast_two_sum(ch, cl2, th, tl1);
    const float tl2 = tl1 + cl1;
    float zh, zl;
    df::fast_two_sum(th, tl2, zh, zl);
    return dfloat(zh, zl);
}

// DWDivDW
__host__ __device__ inline dfloat operator/(dfloat x, dfloat y)
{
    const float th = x.hi / y.hi;
    const dfloat r = times_fp(y, th);
    const float ph = x.hi - r.hi;
    const float dl = x.lo - r.lo;
    const float d = ph + dl;
    const float tl = d / y.hi;
    float zh, zl;
    df::fast_two_sum(th, tl, zh, zl);
    return dfloat(zh, zl);
}

}  // namespace psv

namespace psv {

// FP32 whose multiplies and adds may fuse into FMA instructions: the library is
// built with contraction off (the FP32 and FP64 kernels round every operation as
// the CPU path does), and this type's operators carry the contraction permission
// themselves, so the kernels instantiated with it let the compiler turn a * b + c
// into one FMA.  Results are FP32 results with fewer roundings, not bit-identical
// to the plain FP32 kernels.  Precision PSVFWI_FP32_FMA.
struct __align__(4) ffloat {
    float v;

    ffloat() = default;
    __host__ __device__ constexpr ffloat(float f) : v(f) {}
    __host__ __device__ constexpr ffloat(double d) : v((float)d) {}
    __host__ __device__ constexpr ffloat(int i) : v((float)i) {}
    __host__ __device__ explicit operator double() const { return (double)v; }
};

#pragma clang fp contract(fast)
__host__ __device__ inline ffloat operator+(ffloat a, ffloat b) { return ffloat(a.v + b.v); }
__host__ __device__ inline ffloat operator-(ffloat a, ffloat b) { return ffloat(a.v - b.v); }
__host__ __device__ inline ffloat operator*(ffloat a, ffloat b) { return ffloat(a.v * b.v); }
__host__ __device__ inline ffloat operator/(ffloat a, ffloat b) { return ffloat(a.v / b.v); }
__host__ __device__ inline ffloat operator-(ffloat a) { return ffloat(-a.v); }
#pragma clang fp contract(off)
__host__ __device__ inline bool operator==(ffloat a, ffloat b) { return a.v == b.v; }
__host__ __device__ inline bool operator!=(ffloat a, ffloat b) { return a.v != b.v; }

}  // namespace psv
