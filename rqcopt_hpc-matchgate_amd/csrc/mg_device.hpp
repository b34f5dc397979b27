// Device-side types shared by the HIP kernels and the host launcher.
//
// Amplitude convention (reference src/gate.c:629-651): qubit 0 is the most
// significant bit of the amplitude index; a gate on wires (i, j) touches the
// quadruple x=|00>, y=|01>, z=|10>, w=|11> (first label = wire i).  Every
// instance is stored in the sorted basis (i < j) with the center block reversed
// when the caller's wires were reversed (reference src/gate.c:588-604).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mg {

template <typename R>
struct alignas(2 * sizeof(R)) cpx {  // full-size alignment: one 64-bit (FP32) or 128-bit (FP64) access
    R re, im;
};

__host__ __device__ inline float rfma(float a, float b, float c) { return fmaf(a, b, c); }
__host__ __device__ inline double rfma(double a, double b, double c) { return fma(a, b, c); }

template <typename R>
__host__ __device__ inline cpx<R> cmul(cpx<R> a, cpx<R> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// a*b + c
template <typename R>
__host__ __device__ inline cpx<R> cfma(cpx<R> a, cpx<R> b, cpx<R> c)
{
    return {rfma(a.re, b.re, rfma(-a.im, b.im, c.re)), rfma(a.re, b.im, rfma(a.im, b.re, c.im))};
}

template <typename R>
__host__ __device__ inline cpx<R> cadd(cpx<R> a, cpx<R> b)
{
    return {a.re + b.re, a.im + b.im};
}

template <typename R>
__host__ __device__ inline cpx<R> csub(cpx<R> a, cpx<R> b)
{
    return {a.re - b.re, a.im - b.im};
}

// Raises a kernel's dynamic-LDS limit to `bytes` the first time it is needed and never lowers it,
// so that a launch recorded in a graph stays valid and a launch inside a capture sets nothing.
hipError_t ensure_lds_limit(const void* kernel, size_t bytes);

// One gate instance of a program, in the sorted basis.
struct Inst {
    int32_t hi;     // bit position of the first (smaller) wire: L-1-i
    int32_t lo;     // bit position of the second wire: L-1-j (lo < hi)
    int32_t param;  // parameter gate whose gradient this instance feeds
    int32_t flip;   // caller's wires were reversed: gradient center block is un-reversed
};

// Coefficient block per instance: [0..7] V, [8..15] V^{-1}, [16..23] Z
// (each: center c0..c3 then corner k0..k3, sorted basis).
constexpr int kCoefPerInst = 24;

// amplitude index of quadruple q with zero bits inserted at lo and hi (lo < hi)
__host__ __device__ inline uint32_t quad_base(uint32_t q, int lo, int hi)
{
    uint32_t a = ((q >> lo) << (lo + 1)) | (q & ((1u << lo) - 1u));
    a = ((a >> hi) << (hi + 1)) | (a & ((1u << hi) - 1u));
    return a;
}

}  // namespace mg
