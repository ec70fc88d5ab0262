// CFL-adaptive time step on the device (cfl_util.f90:73-111, main.f90:135-144,
// initialize.f90:191-200, sgs_stag_util.f90:74-82, lagrange_Sdep.f90:430).
//
// With use_cfl_dt every member carries its own clock of CK_N doubles in
// device memory.  A step starts with a reduction of max|u|, max|v|, max|w|
// over i < nx, levels 1..nz-1 and a one-wave update of the clock; every
// kernel that needs dt, tadv1, tadv2, lagran_dt or the pressure scale reads
// its member's slot, so a captured step replays with the dt of the moment and
// nothing waits on the host.
#pragma once

#include <cfloat>

#include "common.hpp"

namespace lg {

enum ClockSlot : int {
    CK_DT = 0,      // dt of the current (last) step
    CK_TADV1,       // AB2 weights
    CK_TADV2,
    CK_LAGDT,       // lagran_dt: displacement time pending for the dynamic model
    CK_TIME,        // total_time
    CK_PSCALE,      // 1 / (nx ny tadv1 dt), press_stag_array.f90:52
    CK_SWAIT,       // statistics: time since the last accumulation
    CK_STIME,       // statistics: accumulated time
    CK_N
};

// max |u|, |v|, |w| of member blockIdx.y over i < nx, levels 1..nz-1 into
// cmax[3 m + c].  Non-negative doubles order as their bit patterns read as
// unsigned integers, so one 64-bit atomic max per block and component merges
// the blocks (cmax starts at 0 = +0.0).
template <typename T>
__global__ void __launch_bounds__(256) k_cfl_max(const T* u, const T* v, const T* w, int nx,
                                                 int ny, int ld, long plane, long mstride, int nz,
                                                 unsigned long long* cmax) {
    const int m = blockIdx.y;
    const long base = (long)m * mstride + plane;   // level 1 of member m
    const long n = (long)(nz - 1) * ny * nx;
    double mu = 0.0, mv = 0.0, mw = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (long)gridDim.x * blockDim.x) {
        const long row = q / nx;                   // (k - 1) * ny + j; plane = ny * ld
        const long o = base + row * ld + (q - row * nx);
        mu = fmax(mu, fabs((double)u[o]));
        mv = fmax(mv, fabs((double)v[o]));
        mw = fmax(mw, fabs((double)w[o]));
    }
    // the block's maxima by wave shuffles (block_reduce), then one 64-bit
    // atomic max per component on the bit patterns of the non-negative
    // doubles
    mu = block_reduce(mu, RedMax{}, 0.0);
    mv = block_reduce(mv, RedMax{}, 0.0);
    mw = block_reduce(mw, RedMax{}, 0.0);
    if (threadIdx.x == 0) {
        atomicMax(cmax + 3 * m, (unsigned long long)__double_as_longlong(mu));
        atomicMax(cmax + 3 * m + 1, (unsigned long long)__double_as_longlong(mv));
        atomicMax(cmax + 3 * m + 2, (unsigned long long)__double_as_longlong(mw));
    }
}

// The head of the time loop for every stepped member: dt_f = dt,
// dt = cfl / max(max|u|/dx, max|v|/dy, max|w|/dz), tadv1 = 1 + dt/(2 dt_f),
// tadv2 = 1 - tadv1; a fresh clock takes dt_f = dt * huge (Euler).  lagran_dt
// grows by dt where lacc[m] (jt >= DYN_init - cs_count + 1 or initu).
// total_time += dt.  Clears the reduction slots for the next step.
static __global__ void k_clock_step(double* clk, unsigned long long* cmax, const int* fresh,
                             const int* lacc, int Ma, double cfl, double dx, double dy, double dz,
                             double cst) {
    for (int m = threadIdx.x; m < Ma; m += blockDim.x) {
        double* c = clk + (long)CK_N * m;
        const double iu = __longlong_as_double((long long)cmax[3 * m]) / dx;
        const double iv = __longlong_as_double((long long)cmax[3 * m + 1]) / dy;
        const double iw = __longlong_as_double((long long)cmax[3 * m + 2]) / dz;
        const double dt = cfl / fmax(iu, fmax(iv, iw));
        const double dt_f = fresh[m] ? dt * DBL_MAX : c[CK_DT];
        const double t1 = 1.0 + (0.5 * dt) / dt_f;
        c[CK_DT] = dt;
        c[CK_TADV1] = t1;
        c[CK_TADV2] = 1.0 - t1;
        if (lacc[m]) c[CK_LAGDT] = c[CK_LAGDT] + dt;
        c[CK_PSCALE] = cst / t1 / dt;
        c[CK_TIME] = c[CK_TIME] + dt;
        cmax[3 * m] = 0ull;
        cmax[3 * m + 1] = 0ull;
        cmax[3 * m + 2] = 0ull;
    }
}

// lagran_dt = 0 after a dynamic-model update (lagrange_Sdep.f90:430).
static __global__ void k_clock_lag_reset(double* clk, const int* mode, int Ma) {
    for (int m = threadIdx.x; m < Ma; m += blockDim.x)
        if (mode[m] > 0) clk[(long)CK_N * m + CK_LAGDT] = 0.0;
}

// Statistics weights: op 0 adds the step's dt to the pending weight, op 1
// drops it (the first sample of a window), op 2 moves it to the accumulated
// time after an accumulation, op 3 clears both (a new window).
static __global__ void k_clock_stats(double* clk, int Ma, int op) {
    for (int m = threadIdx.x; m < Ma; m += blockDim.x) {
        double* c = clk + (long)CK_N * m;
        if (op == 0) {
            c[CK_SWAIT] = c[CK_SWAIT] + c[CK_DT];
        } else {
            if (op == 2) c[CK_STIME] = c[CK_STIME] + c[CK_SWAIT];
            if (op == 3) c[CK_STIME] = 0.0;
            c[CK_SWAIT] = 0.0;
        }
    }
}

}  // namespace lg
