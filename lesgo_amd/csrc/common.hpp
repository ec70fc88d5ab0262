// Shared definitions for the lesgo_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace lg {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }

constexpr double BOGUS = -1234567890.0;   // param.f90:93

// Error status returned through the C ABI (0 = success).
enum Status : int {
    LG_OK = 0,
    LG_ERR_ARG = 1,
    LG_ERR_HIP = 2,
    LG_ERR_ALLOC = 3,
    LG_ERR_STATE = 4,
    LG_ERR_UNSUPPORTED = 5,
    LG_ERR_PIVOT = 6,
};

void set_error(const std::string& msg);

#define LG_HIP(call)                                                         \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            ::lg::set_error(std::string(#call " failed: ") +                 \
                            hipGetErrorString(e_) + " at " + __FILE__ + ":" + \
                            std::to_string(__LINE__));                       \
            throw ::lg::HipError();                                          \
        }                                                                    \
    } while (0)

struct HipError {};
struct ArgError {};
struct PivotError {};

// Reductions over a 64-lane wave by cross-lane shuffles (no LDS): after the
// five halving steps lane 0 holds the result.
struct RedSum {
    __device__ double operator()(double a, double b) const { return a + b; }
};
struct RedMax {
    __device__ double operator()(double a, double b) const { return fmax(a, b); }
};

template <typename Op>
__device__ __forceinline__ double wave_reduce(double v, Op op) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
    return v;
}

// Block reduction (at most 1024 threads): each wave reduces by shuffles,
// the waves' partials meet in LDS and wave 0 reduces them by shuffles; the
// result is in thread 0.  The order is fixed by the block shape, so a
// member's sums are the same alone and in a batch.  Ends with a barrier, so
// the call can be repeated in one kernel.
template <typename Op>
__device__ double block_reduce(double v, Op op, double identity) {
    __shared__ double part[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    v = wave_reduce(v, op);
    if (lane == 0) part[wid] = v;
    __syncthreads();
    double r = identity;
    if (wid == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        r = wave_reduce(lane < nw ? part[lane] : identity, op);
    }
    __syncthreads();
    return r;
}

}  // namespace lg
