// Running time averages (time_average.f90, tavg%compute 176-319, driven as in
// io.f90:558-582) and streamwise spectra, accumulated in FP64 on the device
// whatever the precision of the case.
#include <vector>

#include "case.hpp"
#include "clock.hpp"
#include "pointwise.hpp"

namespace lg {

static constexpr int SBLK = 256;

// the quantities of tavg_t, in this order
enum TavgQ : int {
    TQ_U, TQ_V, TQ_W, TQ_W_UV, TQ_U_W, TQ_V_W, TQ_U2, TQ_V2, TQ_W2, TQ_UV, TQ_UW, TQ_VW,
    TQ_TXX, TQ_TYY, TQ_TZZ, TQ_TXY, TQ_TXZ, TQ_TYZ, TQ_P, TQ_FX, TQ_FY, TQ_FZ, TQ_CS,
    TQ_VORTX, TQ_VORTY, TQ_VORTZ, TQ_N
};

template <typename T>
struct TavgArgs {
    const T *u, *v, *w, *p, *txx, *tyy, *tzz, *txy, *txz, *tyz, *fx, *fy, *fz, *cs;
    const T *dudy, *dudz, *dvdx, *dvdz, *dwdx, *dwdy;
    double* acc;       // quantity q of member m at q * qstride + m * pm
    long qstride, pm;
    double dt;
    const double* clk;     // per-member weights (CFL-adaptive dt) or null: dt
};

// One point (i < nx, j, k = 0..nz) of one member.  Interpolations as
// functions.f90:51-185 (interp_to_uv_grid: var(k+1), var(k) average on
// 1..nz-1, copied to nz; interp_to_w_grid: var(k-1), var(k) on 1..nz), the
// wall rows u_w(1) = v_w(1) = vortz(1) = 0; levels the reference leaves
// unassigned (k = 0 of w_uv, u_w, v_w, vortz) are 0.
template <typename T>
__global__ void k_tavg(TavgArgs<T> a, Geo g) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.nx, m, k, j, i, mb)) return;
    const long P = g.plane;
    const long o = mb + (long)k * P + (long)j * g.ld + i;
    const int nz = g.nz;
    const double dt = a.clk ? a.clk[(long)CK_N * m + CK_SWAIT] : a.dt;
    double* acc = a.acc + (long)m * a.pm + ((long)k * g.ny + j) * g.nx + i;
    auto add = [&](int q, double val) { acc[(long)q * a.qstride] += val * dt; };
    const double u = a.u[o], v = a.v[o], w = a.w[o];
    double wuv = 0.0, fzuv = 0.0;
    if (k >= 1 && k <= nz - 1) {
        wuv = 0.5 * ((double)a.w[o + P] + w);
        fzuv = 0.5 * ((double)a.fz[o + P] + (double)a.fz[o]);
    } else if (k == nz) {
        wuv = 0.5 * (w + (double)a.w[o - P]);
        fzuv = 0.5 * ((double)a.fz[o] + (double)a.fz[o - P]);
    }
    double uw = 0.0, vw = 0.0, vortz = 0.0;
    if (k >= 2) {
        uw = 0.5 * ((double)a.u[o - P] + u);
        vw = 0.5 * ((double)a.v[o - P] + v);
        const double vt0 = (double)a.dvdx[o - P] - (double)a.dudy[o - P];
        const double vt1 = (double)a.dvdx[o] - (double)a.dudy[o];
        vortz = 0.5 * (vt0 + vt1);
    }
    add(TQ_U, u);
    add(TQ_V, v);
    add(TQ_W, w);
    add(TQ_W_UV, wuv);
    add(TQ_U_W, uw);
    add(TQ_V_W, vw);
    add(TQ_U2, u * u);
    add(TQ_V2, v * v);
    add(TQ_W2, w * w);
    add(TQ_UV, u * v);
    add(TQ_UW, uw * w);
    add(TQ_VW, vw * w);
    add(TQ_TXX, (double)a.txx[o]);
    add(TQ_TYY, (double)a.tyy[o]);
    add(TQ_TZZ, (double)a.tzz[o]);
    add(TQ_TXY, (double)a.txy[o]);
    add(TQ_TXZ, (double)a.txz[o]);
    add(TQ_TYZ, (double)a.tyz[o]);
    add(TQ_P, (double)a.p[o] - 0.5 * (u * u + wuv * wuv + v * v));
    if (k >= 1) {
        add(TQ_FX, (double)a.fx[o]);
        add(TQ_FY, (double)a.fy[o]);
        add(TQ_FZ, fzuv);
        add(TQ_CS, (double)a.cs[o]);
    }
    add(TQ_VORTX, (double)a.dwdy[o] - (double)a.dvdz[o]);
    add(TQ_VORTY, (double)a.dudz[o] - (double)a.dwdx[o]);
    add(TQ_VORTZ, vortz);
}

// y-summed spectral power of levels 1..nz-1: grid (kx blocks, (nz-1) * M, 3)
template <typename T>
__global__ void k_spec_acc(const cplx<T>* su, const cplx<T>* sv, const cplx<T>* sw, long msp,
                           long splane, int ny, int NK, int nz, int M, double* acc, double dt,
                           const double* clk) {
    const int kx = blockIdx.x * blockDim.x + threadIdx.x;
    if (kx >= NK) return;
    const int m = blockIdx.y / (nz - 1);
    const int k = blockIdx.y - m * (nz - 1) + 1;
    const int c = blockIdx.z;
    const cplx<T>* s = (c == 0 ? su : (c == 1 ? sv : sw)) + (long)m * msp + (long)k * splane;
    if (clk) dt = clk[(long)CK_N * m + CK_SWAIT];
    double sum = 0.0;
    for (int ky = 0; ky < ny; ++ky) {
        const cplx<T> z = s[(long)ky * NK + kx];
        sum += (double)z.x * (double)z.x + (double)z.y * (double)z.y;
    }
    acc[(((long)c * M + m) * (nz + 1) + k) * NK + kx] += dt * sum;
}

template <typename T>
void Case<T>::stats_enable(int nstart, int nend, int nskip) {
    LG_HIP(hipStreamSynchronize(stream));
    const long pm = (long)(d.nz + 1) * d.ny * d.nx;
    if (!tavg) {
        LG_HIP(hipMalloc(&tavg, (size_t)TQ_N * M * pm * sizeof(double)));
        LG_HIP(hipMalloc(&tspec, (size_t)3 * M * (d.nz + 1) * d.NK * sizeof(double)));
    }
    LG_HIP(hipMemset(tavg, 0, (size_t)TQ_N * M * pm * sizeof(double)));
    LG_HIP(hipMemset(tspec, 0, (size_t)3 * M * (d.nz + 1) * d.NK * sizeof(double)));
    if (cfl_mode)
        hipLaunchKernelGGL(k_clock_stats, dim3(1), dim3(64), 0, stream, d_clock, M, 3);
    st = StatsCfg{};
    st.nstart = nstart;
    st.nend = nend;
    st.nskip = nskip;
    st.on = true;
}

template <typename T>
void Case<T>::stats_after_step() {
    if (jt_total < st.nstart || jt_total > st.nend) return;
    // with use_cfl_dt the weights are each member's, kept in its clock
    if (cfl_mode)
        hipLaunchKernelGGL(k_clock_stats, dim3(1), dim3(64), 0, stream, d_clock, Ma, 0);
    st.dt += d.prm.dt;
    if ((jt_total - st.nstart) % st.nskip != 0) return;
    if (!st.initialized) {
        st.initialized = true;
        st.dt = 0.0;
        if (cfl_mode)
            hipLaunchKernelGGL(k_clock_stats, dim3(1), dim3(64), 0, stream, d_clock, Ma, 1);
        return;
    }
    materialize(SG_UVW | SG_P);
    const long pm = (long)(d.nz + 1) * d.ny * d.nx;
    TavgArgs<T> a{fld[LG_U],    fld[LG_V],    fld[LG_W],    fld[LG_P],    fld[LG_TXX],
                  fld[LG_TYY],  fld[LG_TZZ],  fld[LG_TXY],  fld[LG_TXZ],  fld[LG_TYZ],
                  fld[LG_FXA],  fld[LG_FYA],  fld[LG_FZA],  fld[LG_CS_OPT2],
                  fld[LG_DUDY], fld[LG_DUDZ], fld[LG_DVDX], fld[LG_DVDZ], fld[LG_DWDX],
                  fld[LG_DWDY], tavg,         (long)M * pm, pm,         st.dt,
                  clk()};
    const Geo g{d.nx, d.ny, d.nz, d.ld, d.plane, 0, d.nz + 1, d.field};
    hipLaunchKernelGGL(k_tavg<T>, dim3((unsigned)((d.nx * d.ny + SBLK - 1) / SBLK), (d.nz + 1) * Ma),
                       dim3(SBLK), 0, stream, a, g);
    // spectra of the end-of-step u, v, w (levels 1..nz-1) in scratch spectra
    const T sc = (T)(1.0 / (d.nx * d.ny));
    const T* src[3] = {fld[LG_U], fld[LG_V], fld[LG_W]};
    const T scs[3] = {sc, sc, sc};
    cplx<T>* S[3] = {spec[0], spec[1], spec[2]};
    fwd2d(3, src, scs, d.plane, 1, d.field, S, d.splane, 1, msp, d.nz - 1);
    hipLaunchKernelGGL(k_spec_acc<T>, dim3((unsigned)((d.NK + 63) / 64), (d.nz - 1) * Ma, 3),
                       dim3(64), 0, stream, (const cplx<T>*)spec[0], (const cplx<T>*)spec[1],
                       (const cplx<T>*)spec[2], msp, d.splane, d.ny, d.NK, d.nz, M, tspec, st.dt,
                       clk());
    if (cfl_mode)
        hipLaunchKernelGGL(k_clock_stats, dim3(1), dim3(64), 0, stream, d_clock, Ma, 2);
    st.total_time += st.dt;
    st.dt = 0.0;
}

// Accumulated averaging time of member m (each member's own with use_cfl_dt).
template <typename T>
double Case<T>::stats_time(int m) {
    if (!cfl_mode) return st.total_time;
    double c[CK_N];
    LG_HIP(hipMemcpyAsync(c, d_clock + (long)CK_N * m, sizeof(c), hipMemcpyDeviceToHost, stream));
    LG_HIP(hipStreamSynchronize(stream));
    return c[CK_STIME];
}

template <typename T>
void Case<T>::stats_mean(int m, int q, double* host) {
    if (!tavg || q < 0 || q >= TQ_N) {
        set_error("statistics not enabled or quantity out of range");
        throw ArgError();
    }
    const long pm = (long)(d.nz + 1) * d.ny * d.nx;
    LG_HIP(hipStreamSynchronize(stream));
    LG_HIP(hipMemcpy(host, tavg + (long)q * M * pm + (long)m * pm, pm * sizeof(double),
                     hipMemcpyDeviceToHost));
    const double tt = stats_time(m);
    for (long i = 0; i < pm; ++i) host[i] = tt > 0 ? host[i] / tt : 0.0;
}

// Raw time-weighted sums of quantity q of member m: the arrays of tavg.out
// (time_average.f90:665-716).
template <typename T>
void Case<T>::stats_sums(int m, int q, double* host) {
    if (!tavg || q < 0 || q >= TQ_N) {
        set_error("statistics not enabled or quantity out of range");
        throw ArgError();
    }
    const long pm = (long)(d.nz + 1) * d.ny * d.nx;
    LG_HIP(hipStreamSynchronize(stream));
    LG_HIP(hipMemcpy(host, tavg + (long)q * M * pm + (long)m * pm, pm * sizeof(double),
                     hipMemcpyDeviceToHost));
}

// Continue the sums of a previous run (tavg%init reading tavg.out,
// time_average.f90:119-166): member m's sums of every quantity (TQ_N arrays
// in quantity order) and the averaging time.  Without use_cfl_dt the
// averaging time is one for every member.  Called after stats_enable and
// before the summation starts; the first summation step then only starts
// the summation, as after tavg%init.
template <typename T>
void Case<T>::stats_restore(int m, double total_time, const double* sums) {
    if (!tavg || st.initialized) {
        set_error(tavg ? "the summation has already started" : "statistics not enabled");
        throw ArgError();
    }
    const long pm = (long)(d.nz + 1) * d.ny * d.nx;
    LG_HIP(hipStreamSynchronize(stream));
    for (int q = 0; q < TQ_N; ++q)
        LG_HIP(hipMemcpy(tavg + (long)q * M * pm + (long)m * pm, sums + (long)q * pm,
                         pm * sizeof(double), hipMemcpyHostToDevice));
    if (cfl_mode)
        LG_HIP(hipMemcpy(d_clock + (long)CK_N * m + CK_STIME, &total_time, sizeof(double),
                         hipMemcpyHostToDevice));
    else
        st.total_time = total_time;
}

template <typename T>
void Case<T>::stats_spectrum(int m, int c, double* host) {
    if (!tspec || c < 0 || c > 2) {
        set_error("statistics not enabled or component out of range");
        throw ArgError();
    }
    const long n = (long)(d.nz + 1) * d.NK;
    LG_HIP(hipStreamSynchronize(stream));
    LG_HIP(hipMemcpy(host, tspec + ((long)c * M + m) * n, n * sizeof(double),
                     hipMemcpyDeviceToHost));
    const double tt = stats_time(m);
    for (long r = 0; r < d.nz + 1; ++r)
        for (int kx = 0; kx < d.NK; ++kx) {
            double& e = host[r * d.NK + kx];
            e = tt > 0 ? e / tt : 0.0;
            if (kx > 0) e *= 2.0;
        }
}

// sum over i < nx, all j, k0 <= k <= k1 of (a - b)^2 and b^2 for one field of
// two cases of the same grid (any precisions)
template <typename TA, typename TB>
__global__ void k_diff_l2(const TA* a, const TB* b, int nx, int ny, int ld, int k0, int nlev,
                          double* part) {
    const long n = (long)nlev * ny * nx;
    double d2 = 0.0, b2 = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (long)gridDim.x * blockDim.x) {
        const long k = q / ((long)ny * nx) + k0;
        const long r = q % ((long)ny * nx);
        const long o = (k * ny + r / nx) * ld + r % nx;
        const double x = (double)a[o], y = (double)b[o];
        d2 += (x - y) * (x - y);
        b2 += y * y;
    }
    d2 = block_reduce(d2, RedSum{}, 0.0);
    b2 = block_reduce(b2, RedSum{}, 0.0);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = d2;
        part[2 * blockIdx.x + 1] = b2;
    }
}

template <typename TA, typename TB>
void diff_l2(Case<TA>& a, int ma, Case<TB>& b, int mb, int field, int k0, int k1, double* num,
             double* den) {
    const Dims& d = a.d;
    if (d.nx != b.d.nx || d.ny != b.d.ny || d.nz != b.d.nz) {
        set_error("diff_l2: cases of different grids");
        throw ArgError();
    }
    LG_HIP(hipStreamSynchronize(a.stream));
    LG_HIP(hipStreamSynchronize(b.stream));
    const int nb = 256;
    hipLaunchKernelGGL((k_diff_l2<TA, TB>), dim3(nb), dim3(SBLK), 0, b.stream,
                       (const TA*)a.mem(field, ma), (const TB*)b.mem(field, mb), d.nx, d.ny, d.ld,
                       k0, k1 - k0 + 1, b.red);
    std::vector<double> h(2 * nb);
    LG_HIP(hipMemcpyAsync(h.data(), b.red, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                          b.stream));
    LG_HIP(hipStreamSynchronize(b.stream));
    double s0 = 0.0, s1 = 0.0;
    for (int q = 0; q < nb; ++q) {
        s0 += h[2 * q];
        s1 += h[2 * q + 1];
    }
    *num = s0;
    *den = s1;
}

#define LG_INST_STATS(T)                                              \
    template void Case<T>::stats_enable(int, int, int);               \
    template void Case<T>::stats_after_step();                        \
    template void Case<T>::stats_mean(int, int, double*);             \
    template void Case<T>::stats_sums(int, int, double*);             \
    template void Case<T>::stats_restore(int, double, const double*); \
    template void Case<T>::stats_spectrum(int, int, double*);
LG_INST_STATS(double)
LG_INST_STATS(float)
#undef LG_INST_STATS

template void diff_l2<double, double>(Case<double>&, int, Case<double>&, int, int, int, int,
                                      double*, double*);
template void diff_l2<double, float>(Case<double>&, int, Case<float>&, int, int, int, int,
                                     double*, double*);
template void diff_l2<float, double>(Case<float>&, int, Case<double>&, int, int, int, int,
                                     double*, double*);
template void diff_l2<float, float>(Case<float>&, int, Case<float>&, int, int, int, int, double*,
                                    double*);

}  // namespace lg
