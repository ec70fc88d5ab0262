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

// Bandwidth probes: 1 = one float4 per thread; 2 = four independent float4
// per thread; 3 = as 2 with non-temporal loads and stores; 4 = read-only
// (sum to one value per block); 5 = write-only.
__global__ void __launch_bounds__(256) k_bw1(const float4* __restrict__ a, float4* __restrict__ b,
                                             long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

typedef float vf4 __attribute__((ext_vector_type(4)));

template <bool NT>
__global__ void __launch_bounds__(256) k_bw2(const float4* __restrict__ a4, float4* __restrict__ b4,
                                             long n) {
    const vf4* a = reinterpret_cast<const vf4*>(a4);
    vf4* b = reinterpret_cast<vf4*>(b4);
    const long base = (long)blockIdx.x * (4 * 256) + threadIdx.x;
    vf4 v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const long i = base + q * 256;
        if (i < n) {
            if (NT) v[q] = __builtin_nontemporal_load(a + i);
            else v[q] = a[i];
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const long i = base + q * 256;
        if (i < n) {
            if (NT) __builtin_nontemporal_store(v[q], b + i);
            else b[i] = v[q];
        }
    }
}

__global__ void __launch_bounds__(256) k_bw_read(const float4* __restrict__ a, float* out, long n) {
    const long base = (long)blockIdx.x * (4 * 256) + threadIdx.x;
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const long i = base + q * 256;
        if (i < n) {
            const float4 v = a[i];
            s += v.x + v.y + v.z + v.w;
        }
    }
    if (s == 12345.678f) out[blockIdx.x] = s;     // keeps the loads
}

__global__ void __launch_bounds__(256) k_bw_write(float4* __restrict__ b, long n) {
    const long base = (long)blockIdx.x * (4 * 256) + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const long i = base + q * 256;
        if (i < n) b[i] = make_float4(1.f, 2.f, 3.f, 4.f);
    }
}
}  // namespace lg

// Times `reps` launches of probe `variant` over nbytes (multiple of 16).
extern "C" int lg_bw_probe(int variant, long nbytes, int reps, double* ms) {
    if (nbytes <= 0 || nbytes % 16 != 0 || reps < 1 || !ms || variant < 1 || variant > 5)
        return LG_ERR_ARG;
    float4 *a = nullptr, *b = nullptr;
    float* o = nullptr;
    const long n = nbytes / 16;
    if (hipMalloc(&a, nbytes) != hipSuccess) return LG_ERR_ALLOC;
    if (hipMalloc(&b, nbytes) != hipSuccess || hipMalloc(&o, 1 << 20) != hipSuccess) {
        (void)hipFree(a);
        return LG_ERR_ALLOC;
    }
    (void)hipMemset(a, 0, nbytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const unsigned g1 = (unsigned)((n + 255) / 256), g4 = (unsigned)((n + 1023) / 1024);
    auto launch = [&] {
        switch (variant) {
            case 1: hipLaunchKernelGGL(lg::k_bw1, dim3(g1), dim3(256), 0, 0, a, b, n); break;
            case 2: hipLaunchKernelGGL(lg::k_bw2<false>, dim3(g4), dim3(256), 0, 0, a, b, n); break;
            case 3: hipLaunchKernelGGL(lg::k_bw2<true>, dim3(g4), dim3(256), 0, 0, a, b, n); break;
            case 4: hipLaunchKernelGGL(lg::k_bw_read, dim3(g4), dim3(256), 0, 0, a, o, n); break;
            default: hipLaunchKernelGGL(lg::k_bw_write, dim3(g4), dim3(256), 0, 0, b, n); break;
        }
    };
    launch();
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, 0);
    const hipError_t st = hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(o);
    return st == hipSuccess ? LG_OK : LG_ERR_HIP;
}

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

// Device facts used to size launches: out[0] LDS bytes per workgroup,
// out[1] LDS bytes per CU, out[2] CUs, out[3] L2 bytes, out[4] max threads
// per block, out[5] warp size, out[6] memory clock kHz, out[7] bus width bits.
extern "C" int lg_device_info(long* out, int n) {
    if (!out || n < 8) return LG_ERR_ARG;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return LG_ERR_HIP;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return LG_ERR_HIP;
    out[0] = (long)p.sharedMemPerBlock;
    out[1] = (long)p.maxSharedMemoryPerMultiProcessor;
    out[2] = p.multiProcessorCount;
    out[3] = p.l2CacheSize;
    out[4] = p.maxThreadsPerBlock;
    out[5] = p.warpSize;
    out[6] = p.memoryClockRate;
    out[7] = p.memoryBusWidth;
    return LG_OK;
}
