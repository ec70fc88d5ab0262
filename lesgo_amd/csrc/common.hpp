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

}  // namespace lg
