// Plain device copy used to calibrate HBM byte counters and to measure the
// achievable copy bandwidth of the card (no counterpart in the reference).
#include <algorithm>
#include "common.hpp"
#include "../../include/lesgo_amd.h"

namespace lg {
__global__ void __launch_bounds__(256) k_copy16(const float4* __restrict__ a, float4* __restrict__ b,
                                                long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        b[i] = a[i];
}
}  // namespace lg

extern "C" int lg_calibrate_copy(long nbytes, int reps, double* ms_per_copy) {
    if (nbytes <= 0 || nbytes % 16 != 0 || reps < 1 || !ms_per_copy) return LG_ERR_ARG;
    float4 *a = nullptr, *b = nullptr;
    hipEvent_t e0, e1;
    const long n = nbytes / 16;
    if (hipMalloc(&a, nbytes) != hipSuccess) return LG_ERR_ALLOC;
    if (hipMalloc(&b, nbytes) != hipSuccess) {
        (void)hipFree(a);
        return LG_ERR_ALLOC;
    }
    (void)hipMemset(a, 0, nbytes);
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    // one 16-byte element per thread over the whole array: many waves in
    // flight (a few blocks per CU looping over it moved ~30 % less here)
    const long blocks = std::min((n + 255) / 256, 2147483647L);
    hipLaunchKernelGGL(lg::k_copy16, dim3((unsigned)blocks), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(lg::k_copy16, dim3((unsigned)blocks), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    hipError_t st = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_copy = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return st == hipSuccess ? LG_OK : LG_ERR_HIP;
}
