// Subsystem drivers: each method replaces one reference subroutine and runs
// on the case's stream.  See the kernel headers for the per-loop citations.
#include <algorithm>
#include <cmath>
#include <vector>

#include "case.hpp"
#include "fft2.hpp"
#include "lagrange.hpp"
#include "pointwise.hpp"
#include "pressure.hpp"

namespace lg {

static constexpr double PI_D = 3.1415926535897932384626433;   // param.f90:94
static constexpr int BLK = 256;
static constexpr size_t LDS_MAX = 64 * 1024;

#define LG_SPLITS(X) X(4, 4) X(4, 6) X(4, 8) X(6, 8) X(8, 8) X(8, 9) X(8, 12) X(8, 16) \
    X(12, 12) X(12, 16) X(16, 16) X(16, 24)

template <typename T>
static bool v2_r2c(int n1, int n2, dim3 g, size_t lds, hipStream_t s, const RowR2CArgs<T>& a) {
#define LG_CASE(A, B)                                                                    \
    if (n1 == A && n2 == B) {                                                            \
        hipLaunchKernelGGL((k_rows_r2c_v2<T, A, B>), g, dim3(BLK), lds, s, a);           \
        return true;                                                                     \
    }
    LG_SPLITS(LG_CASE)
#undef LG_CASE
    return false;
}

template <typename T>
static bool v2_c2r(int n1, int n2, dim3 g, size_t lds, hipStream_t s, const RowC2RArgs<T>& a) {
#define LG_CASE(A, B)                                                                    \
    if (n1 == A && n2 == B) {                                                            \
        hipLaunchKernelGGL((k_rows_c2r_v2<T, A, B>), g, dim3(BLK), lds, s, a);           \
        return true;                                                                     \
    }
    LG_SPLITS(LG_CASE)
#undef LG_CASE
    return false;
}

template <typename T, int DIR>
static bool v2_cols(int n1, int n2, dim3 g, size_t lds, hipStream_t s, const ColArgs<T>& a) {
#define LG_CASE(A, B)                                                                    \
    if (n1 == A && n2 == B) {                                                            \
        hipLaunchKernelGGL((k_cols_v2<T, A, B, DIR>), g, dim3(BLK), lds, s, a);          \
        return true;                                                                     \
    }
    LG_SPLITS(LG_CASE)
#undef LG_CASE
    return false;
}

// Launch a column pass: the register four-step kernel when the length has a
// split, the generic LDS kernel otherwise.  Sets c.TC / c.stride.
template <typename T, int DIR>
static void launch_cols(ColArgs<T>& c, int njobs, int nlev, hipStream_t s) {
    const int N = c.plan.N;
    int n1, n2;
    if (split_of(N, n1, n2)) {
        const int YS = n1 * (n2 + 1) + 1;
        int TC = 16;
        while (TC > 1 && (size_t)TC * YS * sizeof(cplx<T>) > LDS_MAX) TC /= 2;
        c.TC = std::min(TC, c.ncols);
        c.stride = YS;
        const size_t lds = (size_t)c.TC * YS * sizeof(cplx<T>);
        dim3 grid(njobs, (c.ncols + c.TC - 1) / c.TC, nlev);
        if (v2_cols<T, DIR>(n1, n2, grid, lds, s, c)) return;
    }
    int TC = 16;
    while (TC > 1 && (size_t)TC * (N + 1) * sizeof(cplx<T>) > LDS_MAX) TC /= 2;
    c.TC = std::min(TC, c.ncols);
    c.stride = N + 1;
    const size_t lds = (size_t)c.TC * c.stride * sizeof(cplx<T>);
    dim3 grid(njobs, (c.ncols + c.TC - 1) / c.TC, nlev);
    hipLaunchKernelGGL((k_cols<T, DIR>), grid, dim3(BLK), lds, s, c);
}

Dims::Dims(const lg_params& p) : prm(p) {
    nx = p.nx;
    ny = p.ny;
    nz = p.Nz + 1;
    lh = nx / 2 + 1;
    ld = 2 * lh;
    nx2 = 3 * nx / 2;
    ny2 = 3 * ny / 2;
    ld_big = 2 * (nx2 / 2 + 1);
    NK = nx / 2;
    dx = p.lx / nx;
    dy = p.ly / ny;
    dz = p.lz / (nz - 1);
    tadv1 = 1.5;
    tadv2 = 1.0 - tadv1;
    delta = 1 * std::pow(dx * dy * dz, 1.0 / 3.0);
    lagran_dt = p.cs_count * p.dt;
    plane = (long)ny * ld;
    field = (long)(nz + 1) * plane;
    splane = (long)ny * NK;
    bplane = (long)ny2 * ld_big;
    bsplane = (long)ny2 * NK;
}

// ---------------------------------------------------------------------------
// FFT tables
// ---------------------------------------------------------------------------
PlanTables::~PlanTables() {
    if (perm) (void)hipFree(perm);
    if (tw64) (void)hipFree(tw64);
    if (tw32) (void)hipFree(tw32);
}

void PlanTables::build(int N) {
    std::vector<int> rs;
    int n = N;
    for (int r : {4, 3, 2}) {
        while (n % r == 0) {
            rs.push_back(r);
            n /= r;
        }
    }
    if (n != 1 || (int)rs.size() > MAX_STAGES) {
        set_error("transform length " + std::to_string(N) + " is not of the form 2^a 3^b");
        throw ArgError();
    }
    const int m = (int)rs.size();
    // slot P holds natural index rev(P): P = sum_s q_s * (N / (r_{m-1}..r_s)),
    // n = q_{m-1} + r_{m-1} (q_{m-2} + r_{m-2} (... + r_1 q_0))
    std::vector<int> h_perm(N);
    for (int P = 0; P < N; ++P) {
        std::vector<int> q(m);
        int rem = P, blk = N;
        for (int s = m - 1; s >= 0; --s) {
            blk /= rs[s];
            q[s] = rem / blk;
            rem %= blk;
        }
        int nn = 0;
        for (int s = 0; s < m; ++s) nn = q[s] + rs[s] * nn;
        h_perm[nn] = P;
    }
    std::vector<double2> h64(N);
    std::vector<float2> h32(N);
    const long double pil = 3.141592653589793238462643383279502884L;
    for (int j = 0; j < N; ++j) {
        const long double ang = 2.0L * pil * (long double)j / (long double)N;
        const double c = (double)cosl(ang);
        const double s = (double)-sinl(ang);
        h64[j] = make_double2(c, s);
        h32[j] = make_float2((float)cosl(ang), (float)-sinl(ang));
    }
    LG_HIP(hipMalloc(&perm, N * sizeof(int)));
    LG_HIP(hipMalloc(&tw64, N * sizeof(double2)));
    LG_HIP(hipMalloc(&tw32, N * sizeof(float2)));
    LG_HIP(hipMemcpy(perm, h_perm.data(), N * sizeof(int), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(tw64, h64.data(), N * sizeof(double2), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(tw32, h32.data(), N * sizeof(float2), hipMemcpyHostToDevice));
    plan.N = N;
    plan.nstage = m;
    for (int s = 0; s < m; ++s) plan.radix[s] = rs[s];
    plan.perm = perm;
    plan.tw64 = tw64;
    plan.tw32 = tw32;
}

TurbineSet::~TurbineSet() {
    if (node_lin) (void)hipFree(node_lin);
    if (node_ind) (void)hipFree(node_ind);
    if (off) (void)hipFree(off);
    if (node_tid) (void)hipFree(node_tid);
    if (coef) (void)hipFree(coef);
    if (state) (void)hipFree(state);
}

// ---------------------------------------------------------------------------
// Construction
// ---------------------------------------------------------------------------
template <typename T>
static T* dalloc(long n) {
    T* p = nullptr;
    LG_HIP(hipMalloc(&p, (size_t)n * sizeof(T)));
    LG_HIP(hipMemset(p, 0, (size_t)n * sizeof(T)));
    return p;
}

template <typename T>
Case<T>::Case(const lg_params& p) : d(p) {
    LG_HIP(hipStreamCreate(&stream));
    for (int i = 0; i < LG_NFIELDS; ++i) fld[i] = dalloc<T>(d.field);
    for (int i = 0; i < NSPEC; ++i) spec[i] = dalloc<cplx<T>>((long)(d.nz + 1) * d.splane);
    for (int i = 0; i < NBIG; ++i) big[i] = dalloc<T>((long)(d.nz + 1) * d.bplane);
    for (int i = 0; i < 6; ++i) bigspec[i] = dalloc<cplx<T>>((long)(d.nz + 1) * d.bsplane);
    for (int i = 0; i < NTMP; ++i) tmp[i] = dalloc<T>(d.field);
    sU = dalloc<cplx<T>>((long)(d.nz + 1) * d.splane);
    sV = dalloc<cplx<T>>((long)(d.nz + 1) * d.splane);
    sW = dalloc<cplx<T>>((long)(d.nz + 1) * d.splane);
    gam = dalloc<double>((long)(d.nz + 1) * d.splane);
    usc = dalloc<double2>((long)(d.nz + 1) * d.splane);
    dev_err = dalloc<int>(1);
    lag_chunk = std::min(d.nz, 64);
    lag_phys = dalloc<T>(15L * lag_chunk * d.plane);
    lag_spec = dalloc<cplx<T>>(21L * lag_chunk * d.splane);
    lag_filt = dalloc<T>(42L * lag_chunk * d.plane);
    Ftmp = dalloc<T>(4L * d.field);
    red = dalloc<double>(1024);
    px.build(d.nx);
    py.build(d.ny);
    pbx.build(d.nx2);
    pby.build(d.ny2);
    // test-filter kernels (test_filtermodule.f90:38-123), ifilter = 1
    std::vector<T> g1((size_t)d.splane), g2((size_t)d.splane);
    const double kxs = 2.0 * PI_D / p.lx, kys = 2.0 * PI_D / p.ly;
    const double kc2a = std::pow(PI_D / (2.0 * std::sqrt(d.dx * d.dy)), 2);
    const double kc2b = std::pow(PI_D / (4.0 * std::sqrt(d.dx * d.dy)), 2);
    const double norm = 1.0 / (d.nx * d.ny);
    for (int j = 0; j < d.ny; ++j) {
        for (int i = 0; i < d.NK; ++i) {
            double kx = kxs * i, ky = kys * ((j < d.ny / 2) ? j : j - d.ny);
            if (j == d.ny / 2) kx = ky = 0.0;
            const double k2 = kx * kx + ky * ky;
            double a = (k2 >= kc2a) ? 0.0 : norm;
            double b = (k2 >= kc2b) ? 0.0 : norm;
            if (j == d.ny / 2) a = b = 0.0;
            g1[(size_t)j * d.NK + i] = (T)a;
            g2[(size_t)j * d.NK + i] = (T)b;
        }
    }
    G1 = dalloc<T>(d.splane);
    G2 = dalloc<T>(d.splane);
    LG_HIP(hipMemcpy(G1, g1.data(), d.splane * sizeof(T), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(G2, g2.data(), d.splane * sizeof(T), hipMemcpyHostToDevice));
}

template <typename T>
Case<T>::~Case() {
    if (stream) (void)hipStreamSynchronize(stream);
    auto fr = [](void* p) {
        if (p) (void)hipFree(p);
    };
    for (auto p : fld) fr(p);
    for (auto p : spec) fr(p);
    for (auto p : big) fr(p);
    for (auto p : bigspec) fr(p);
    for (auto p : tmp) fr(p);
    fr(sU);
    fr(sV);
    fr(sW);
    fr(gam);
    fr(usc);
    fr(dev_err);
    fr(lag_phys);
    fr(lag_spec);
    fr(lag_filt);
    fr(Ftmp);
    fr(red);
    fr(G1);
    fr(G2);
    if (stream) (void)hipStreamDestroy(stream);
}

// ---------------------------------------------------------------------------
// Transform helpers
// ---------------------------------------------------------------------------
template <typename T>
static int rows_per_wg(int N, int npairs) {
    int S = std::max(1, 2048 / N);
    while (S > 1 && (size_t)S * (N + 1) * sizeof(cplx<T>) > LDS_MAX) --S;
    return std::min(S, npairs);
}

template <typename T>
static int cols_per_tile(int N, int ncols) {
    int TC = 16;
    while (TC > 1 && (size_t)TC * (N + 1) * sizeof(cplx<T>) > LDS_MAX) TC /= 2;
    return std::min(TC, ncols);
}

template <typename T>
static void launch_rows_r2c(KernelTimer& tm, const FftPlan& pl, int njobs, const T* const* src,
                            const T* scales, long src_plane, int src_ld, int src_lev0,
                            cplx<T>* const* dst, long dst_plane, int dst_lev0, int nrows,
                            int ncols, int nlev, hipStream_t s) {
    for (int j0 = 0; j0 < njobs; j0 += MAX_JOBS) {
        const int nj = std::min(MAX_JOBS, njobs - j0);
        RowR2CArgs<T> a{};
        a.plan = pl;
        for (int i = 0; i < nj; ++i) {
            a.src[i] = src[j0 + i];
            a.dst[i] = dst[j0 + i];
            a.scale[i] = scales ? scales[j0 + i] : (T)1;
        }
        a.src_plane = src_plane;
        a.dst_plane = dst_plane;
        a.src_ld = src_ld;
        a.nrows = nrows;
        a.ncols = ncols;
        a.lev0_src = src_lev0;
        a.lev0_dst = dst_lev0;
        int n1, n2;
        const bool v2 = split_of(pl.N, n1, n2);
        size_t lds;
        if (v2) {
            const int YS = n1 * (n2 + 1) + 1;
            int S = std::min(std::min(BLK / n1, 16), nrows / 2);
            while (S > 1 && (size_t)S * YS * sizeof(cplx<T>) > LDS_MAX) --S;
            a.S = S;
            a.stride = YS;
            lds = (size_t)a.S * YS * sizeof(cplx<T>);
        } else {
            a.stride = pl.N + 1;
            a.S = rows_per_wg<T>(pl.N, nrows / 2);
            lds = (size_t)a.S * a.stride * sizeof(cplx<T>);
        }
        dim3 grid((nrows / 2 + a.S - 1) / a.S, nlev, nj);
        const double bytes = (double)nj * nlev * nrows * ((double)pl.N * sizeof(T) +
                                                           (double)ncols * sizeof(cplx<T>));
        tm.run(s, KK_ROWS_R2C, bytes, [&] {
            if (!(v2 && v2_r2c<T>(n1, n2, grid, lds, s, a)))
                hipLaunchKernelGGL(k_rows_r2c<T>, grid, dim3(BLK), lds, s, a);
        });
    }
}

template <typename T>
static void launch_rows_c2r(KernelTimer& tm, const FftPlan& pl, int njobs,
                            const cplx<T>* const* src, long src_plane, int src_lev0,
                            T* const* dst, long dst_plane, int dst_ld, int dst_lev0, int nrows,
                            int ncols, int nlev, hipStream_t s) {
    for (int j0 = 0; j0 < njobs; j0 += MAX_JOBS) {
        const int nj = std::min(MAX_JOBS, njobs - j0);
        RowC2RArgs<T> a{};
        a.plan = pl;
        for (int i = 0; i < nj; ++i) {
            a.src[i] = src[j0 + i];
            a.dst[i] = dst[j0 + i];
        }
        a.src_plane = src_plane;
        a.dst_plane = dst_plane;
        a.dst_ld = dst_ld;
        a.nrows = nrows;
        a.ncols = ncols;
        a.lev0_src = src_lev0;
        a.lev0_dst = dst_lev0;
        int n1, n2;
        const bool v2 = split_of(pl.N, n1, n2);
        size_t lds;
        if (v2) {
            const int YS = n1 * (n2 + 1) + 1;
            int S = std::min(16, nrows / 2);
            while (S > 1 && (size_t)S * YS * sizeof(cplx<T>) > LDS_MAX) --S;
            a.S = S;
            a.stride = YS;
            lds = (size_t)a.S * YS * sizeof(cplx<T>);
        } else {
            a.stride = pl.N + 1;
            a.S = rows_per_wg<T>(pl.N, nrows / 2);
            lds = (size_t)a.S * a.stride * sizeof(cplx<T>);
        }
        dim3 grid((nrows / 2 + a.S - 1) / a.S, nlev, nj);
        const double bytes = (double)nj * nlev * nrows * ((double)ncols * sizeof(cplx<T>) +
                                                           (double)dst_ld * sizeof(T));
        tm.run(s, KK_ROWS_C2R, bytes, [&] {
            if (!(v2 && v2_c2r<T>(n1, n2, grid, lds, s, a)))
                hipLaunchKernelGGL(k_rows_c2r<T>, grid, dim3(BLK), lds, s, a);
        });
    }
}

template <typename T>
void Case<T>::fwd2d(int njobs, const T* const* src, const T* scales, long src_plane, int src_lev0,
                    cplx<T>* const* dst, long dst_plane, int dst_lev0, int nlev) {
    if (nlev <= 0) return;
    launch_rows_r2c<T>(timer, px.plan, njobs, src, scales, src_plane, d.ld, src_lev0, dst, dst_plane,
                       dst_lev0, d.ny, d.NK, nlev, stream);
    for (int j0 = 0; j0 < njobs; j0 += MAX_JOBS) {
        const int nj = std::min(MAX_JOBS, njobs - j0);
        ColArgs<T> c{};
        c.plan = py.plan;
        for (int i = 0; i < nj; ++i) {
            c.src[i] = dst[j0 + i];
            c.dst[i] = dst[j0 + i];
            c.op[i] = OP_NONE;
        }
        c.src_plane = c.dst_plane = dst_plane;
        c.lev0_src = c.lev0_dst = dst_lev0;
        c.ncols = d.NK;
        c.ny = d.ny;
        c.map = MAP_ID;
        c.zero_nyq = 1;
        c.TC = cols_per_tile<T>(d.ny, d.NK);
        c.stride = d.ny + 1;
        c.kxs = (T)(2.0 * PI_D / d.prm.lx);
        c.kys = (T)(2.0 * PI_D / d.prm.ly);
        c.G1 = G1;
        c.G2 = G2;
        const size_t lds = (size_t)c.TC * c.stride * sizeof(cplx<T>);
        dim3 grid(nj, (d.NK + c.TC - 1) / c.TC, nlev);
        const int nj_ = nj, nlev_ = nlev;
        const double cbytes = (double)nj_ * nlev_ * c.ncols * sizeof(cplx<T>) * (d.ny + d.ny);
        timer.run(stream, KK_COLS_FWD, cbytes, [&] { launch_cols<T, -1>(c, nj_, nlev_, stream); });
    }
}

template <typename T>
void Case<T>::inv2d(int njobs, const cplx<T>* const* src, const int* ops, long src_plane,
                    int src_lev0, T* const* dst, long dst_plane, int dst_lev0, int nlev,
                    cplx<T>* const* tmpspec, long tmp_plane, const cplx<T>* const* src2) {
    if (nlev <= 0) return;
    for (int j0 = 0; j0 < njobs; j0 += MAX_JOBS) {
        const int nj = std::min(MAX_JOBS, njobs - j0);
        ColArgs<T> c{};
        c.plan = py.plan;
        for (int i = 0; i < nj; ++i) {
            c.src[i] = src[j0 + i];
            c.src2[i] = src2 ? src2[j0 + i] : nullptr;
            c.dst[i] = tmpspec[i];
            c.op[i] = ops[j0 + i];
        }
        c.src_plane = src_plane;
        c.dst_plane = tmp_plane;
        c.lev0_src = src_lev0;
        c.lev0_dst = 0;
        c.ncols = d.NK;
        c.ny = d.ny;
        c.map = MAP_ID;
        c.zero_nyq = 0;
        c.TC = cols_per_tile<T>(d.ny, d.NK);
        c.stride = d.ny + 1;
        c.kxs = (T)(2.0 * PI_D / d.prm.lx);
        c.kys = (T)(2.0 * PI_D / d.prm.ly);
        c.G1 = G1;
        c.G2 = G2;
        const size_t lds = (size_t)c.TC * c.stride * sizeof(cplx<T>);
        dim3 grid(nj, (d.NK + c.TC - 1) / c.TC, nlev);
        const int nj_ = nj, nlev_ = nlev;
        const double cbytes = (double)nj_ * nlev_ * c.ncols * sizeof(cplx<T>) * (d.ny + d.ny);
        timer.run(stream, KK_COLS_INV, cbytes, [&] { launch_cols<T, +1>(c, nj_, nlev_, stream); });
        std::vector<const cplx<T>*> ts(tmpspec, tmpspec + nj);
        launch_rows_c2r<T>(timer, px.plan, nj, ts.data(), tmp_plane, 0, dst + j0, dst_plane, d.ld,
                           dst_lev0, d.ny, d.NK, nlev, stream);
    }
}

template <typename T>
void Case<T>::pad_inv_big(int njobs, const cplx<T>* const* src, int lev0, int nlev) {
    ColArgs<T> c{};
    c.plan = pby.plan;
    for (int i = 0; i < njobs; ++i) {
        c.src[i] = src[i];
        c.dst[i] = bigspec[i];
        c.op[i] = OP_NONE;
    }
    c.src_plane = d.splane;
    c.dst_plane = d.bsplane;
    c.lev0_src = lev0;
    c.lev0_dst = lev0;
    c.ncols = d.NK;
    c.ny = d.ny;
    c.map = MAP_PAD;
    c.zero_nyq = 0;
    c.TC = cols_per_tile<T>(d.ny2, d.NK);
    c.stride = d.ny2 + 1;
    c.G1 = G1;
    c.G2 = G2;
    const size_t lds = (size_t)c.TC * c.stride * sizeof(cplx<T>);
    dim3 grid(njobs, (d.NK + c.TC - 1) / c.TC, nlev);
    const int nj_ = njobs, nlev_ = nlev;
        const double cbytes = (double)nj_ * nlev_ * c.ncols * sizeof(cplx<T>) * (d.ny + d.ny2);
        timer.run(stream, KK_COLS_INV, cbytes, [&] { launch_cols<T, +1>(c, nj_, nlev_, stream); });
    std::vector<const cplx<T>*> bs(bigspec, bigspec + njobs);
    launch_rows_c2r<T>(timer, pbx.plan, njobs, bs.data(), d.bsplane, lev0, big, d.bplane, d.ld_big, lev0,
                       d.ny2, d.NK, nlev, stream);
}

template <typename T>
void Case<T>::fwd_big_unpad(int njobs, const T* const* src, cplx<T>* const* dst, int lev0,
                            int nlev) {
    launch_rows_r2c<T>(timer, pbx.plan, njobs, src, nullptr, d.bplane, d.ld_big, lev0, bigspec,
                       d.bsplane, lev0, d.ny2, d.NK, nlev, stream);
    ColArgs<T> c{};
    c.plan = pby.plan;
    for (int i = 0; i < njobs; ++i) {
        c.src[i] = bigspec[i];
        c.dst[i] = dst[i];
        c.op[i] = OP_NONE;
    }
    c.src_plane = d.bsplane;
    c.dst_plane = d.splane;
    c.lev0_src = lev0;
    c.lev0_dst = lev0;
    c.ncols = d.NK;
    c.ny = d.ny;
    c.map = MAP_UNPAD;
    c.zero_nyq = 1;
    c.TC = cols_per_tile<T>(d.ny2, d.NK);
    c.stride = d.ny2 + 1;
    c.G1 = G1;
    c.G2 = G2;
    const size_t lds = (size_t)c.TC * c.stride * sizeof(cplx<T>);
    dim3 grid(njobs, (d.NK + c.TC - 1) / c.TC, nlev);
    const int nj_ = njobs, nlev_ = nlev;
        const double cbytes = (double)nj_ * nlev_ * c.ncols * sizeof(cplx<T>) * (d.ny2 + d.ny);
        timer.run(stream, KK_COLS_FWD, cbytes, [&] { launch_cols<T, -1>(c, nj_, nlev_, stream); });
}

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
static inline Geo geo_of(const Dims& d) { return Geo{d.nx, d.ny, d.nz, d.ld, d.plane}; }

static inline dim3 grid_ij(int W, int ny, int nlev) {
    return dim3((unsigned)((W * ny + BLK - 1) / BLK), (unsigned)nlev);
}

template <typename T>
void Case<T>::fill_level(int id, int k, double v) {
    T* p = fld[id] + (long)k * d.plane;
    hipLaunchKernelGGL(k_fill_planes<T>, dim3((unsigned)((d.plane + BLK - 1) / BLK)), dim3(BLK),
                       0, stream, p, d.plane, d.plane, (T)v);
}

template <typename T>
static void fill_ptr(T* p, long n, double v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_planes<T>, dim3((unsigned)((n + BLK - 1) / BLK)), dim3(BLK), 0, s,
                       p, n, n, (T)v);
}

// ---------------------------------------------------------------------------
// derivatives.f90
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::filt_da(int f, int dfdx, int dfdy, cplx<T>* keep) {
    const T sc = (T)(1.0 / (d.nx * d.ny));
    const T* src[1] = {fld[f]};
    cplx<T>* sp = keep ? keep : spec[0];
    cplx<T>* S0[1] = {sp};
    fwd2d(1, src, &sc, d.plane, 0, S0, d.splane, 0, d.nz + 1);
    const cplx<T>* s3[3] = {sp, sp, sp};
    const int ops[3] = {OP_NONE, OP_IKX, OP_IKY};
    T* dst[3] = {fld[f], fld[dfdx], fld[dfdy]};
    cplx<T>* tsp[3] = {spec[1], spec[2], spec[3]};
    inv2d(3, s3, ops, d.splane, 0, dst, d.plane, 0, d.nz + 1, tsp, d.splane);
}

template <typename T>
void Case<T>::ddxy(int f, int dfdx, int dfdy) {
    const T sc = (T)(1.0 / (d.nx * d.ny));
    const T* src[1] = {fld[f]};
    cplx<T>* S0[1] = {spec[0]};
    fwd2d(1, src, &sc, d.plane, 0, S0, d.splane, 0, d.nz + 1);
    const cplx<T>* s2[2] = {spec[0], spec[0]};
    int ops[2] = {OP_IKX, OP_IKY};
    T* dst[2] = {dfdx >= 0 ? fld[dfdx] : nullptr, dfdy >= 0 ? fld[dfdy] : nullptr};
    cplx<T>* tsp[2] = {spec[1], spec[2]};
    if (dfdx < 0) {
        s2[0] = spec[0];
        ops[0] = OP_IKY;
        dst[0] = fld[dfdy];
        inv2d(1, s2, ops, d.splane, 0, dst, d.plane, 0, d.nz + 1, tsp, d.splane);
    } else if (dfdy < 0) {
        inv2d(1, s2, ops, d.splane, 0, dst, d.plane, 0, d.nz + 1, tsp, d.splane);
    } else {
        inv2d(2, s2, ops, d.splane, 0, dst, d.plane, 0, d.nz + 1, tsp, d.splane);
    }
}

template <typename T>
void Case<T>::ddx(int f, int dfdx) { ddxy(f, dfdx, -1); }

template <typename T>
void Case<T>::ddy(int f, int dfdy) { ddxy(f, -1, dfdy); }

template <typename T>
void Case<T>::ddz_uv(int f, int dfdz) {
    const Geo g = geo_of(d);
    hipLaunchKernelGGL(k_ddz_uv<T>, grid_ij(d.nx, d.ny, d.nz), dim3(BLK), 0, stream, fld[f],
                       fld[dfdz], g, (T)(1.0 / d.dz));
    fill_level(dfdz, 0, BOGUS);
    fill_level(dfdz, 1, BOGUS);
    fill_level(dfdz, d.nz, BOGUS);
}

template <typename T>
void Case<T>::ddz_w(int f, int dfdz) {
    const Geo g = geo_of(d);
    hipLaunchKernelGGL(k_ddz_w<T>, grid_ij(d.nx, d.ny, d.nz), dim3(BLK), 0, stream, fld[f],
                       fld[dfdz], g, (T)(1.0 / d.dz));
    fill_level(dfdz, 0, BOGUS);
    fill_level(dfdz, d.nz, BOGUS);
}

template <typename T>
void Case<T>::test_filter_plane(const T* in, T* out, int op) {
    const T* src[1] = {in};
    const T one = (T)1;
    cplx<T>* S0[1] = {spec[4]};
    fwd2d(1, src, &one, d.plane, 0, S0, d.splane, 0, 1);
    const cplx<T>* s1[1] = {spec[4]};
    T* dst[1] = {out};
    cplx<T>* tsp[1] = {spec[5]};
    inv2d(1, s1, &op, d.splane, 0, dst, d.plane, 0, 1, tsp, d.splane);
}

// ---------------------------------------------------------------------------
// wallstress.f90 (lbc_mom = 2, ubc_mom = 0)
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::wallstress() {
    const T one = (T)1;
    const T* src[2] = {lev(LG_U, 1), lev(LG_V, 1)};
    const T sc[2] = {one, one};
    cplx<T>* S0[2] = {spec[4], spec[5]};
    fwd2d(2, src, sc, d.plane, 0, S0, d.splane, 0, 1);
    const cplx<T>* s2[2] = {spec[4], spec[5]};
    const int ops[2] = {OP_G1, OP_G1};
    T* dst[2] = {tmp[0], tmp[1]};
    cplx<T>* tsp[2] = {spec[6], spec[7]};
    inv2d(2, s2, ops, d.splane, 0, dst, d.plane, 0, 1, tsp, d.splane);
    const Geo g = geo_of(d);
    hipLaunchKernelGGL(k_wall_lbc<T>, grid_ij(d.nx, d.ny, 1), dim3(BLK), 0, stream, tmp[0],
                       tmp[1], lev(LG_U, 1), lev(LG_V, 1), lev(LG_TXZ, 1), lev(LG_TYZ, 1),
                       lev(LG_DUDZ, 1), lev(LG_DVDZ, 1), g, std::log(0.5 * d.dz / d.prm.zo),
                       d.prm.vonk, d.dz);
    fill_level(LG_TXZ, d.nz, 0.0);
    fill_level(LG_TYZ, d.nz, 0.0);
    fill_level(LG_DUDZ, d.nz, 0.0);
    fill_level(LG_DVDZ, d.nz, 0.0);
}

// ---------------------------------------------------------------------------
// sgs_stag_util.f90
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::calc_sij() {
    SijArgs<T> a{fld[LG_DUDX], fld[LG_DUDY], fld[LG_DUDZ], fld[LG_DVDX], fld[LG_DVDY],
                 fld[LG_DVDZ], fld[LG_DWDX], fld[LG_DWDY], fld[LG_DWDZ], fld[LG_S11],
                 fld[LG_S12], fld[LG_S13], fld[LG_S22], fld[LG_S23], fld[LG_S33]};
    hipLaunchKernelGGL(k_calc_sij<T>, grid_ij(d.nx, d.ny, d.nz), dim3(BLK), 0, stream, a,
                       geo_of(d));
}

template <typename T>
void Case<T>::sgs_stag() {
    calc_sij();
    if (jt_total % d.prm.cs_count == 0) lagrange_sdep();
    TauArgs<T> a{fld[LG_S11], fld[LG_S12], fld[LG_S13], fld[LG_S22], fld[LG_S23], fld[LG_S33],
                 fld[LG_CS_OPT2], fld[LG_NU_T], fld[LG_TXX], fld[LG_TXY], fld[LG_TYY],
                 fld[LG_TXZ], fld[LG_TYZ], fld[LG_TZZ], (T)(d.delta * d.delta), (T)d.prm.nu};
    hipLaunchKernelGGL(k_sgs_tau<T>, grid_ij(d.nx, d.ny, d.nz), dim3(BLK), 0, stream, a,
                       geo_of(d));
    for (int id : {LG_TXX, LG_TXY, LG_TXZ, LG_TYY, LG_TYZ, LG_TZZ}) fill_level(id, 0, BOGUS);
    for (int id : {LG_TXX, LG_TXY, LG_TYY, LG_TZZ}) fill_level(id, d.nz, BOGUS);
}

// ---------------------------------------------------------------------------
// interpolag_Sdep.f90 / lagrange_Sdep.f90
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::interpolag_sdep() {
    const int ids[4] = {LG_F_LM, LG_F_MM, LG_F_QN, LG_F_NN};
    for (int q = 0; q < 4; ++q)
        LG_HIP(hipMemcpyAsync(Ftmp + q * d.field, fld[ids[q]], d.field * sizeof(T),
                              hipMemcpyDeviceToDevice, stream));
    InterpGeo g{d.nx, d.ny, d.nz, d.ld, d.plane, d.dx, d.dy, d.dz,
                d.prm.lx, d.prm.ly, d.prm.lz, d.lagran_dt};
    hipLaunchKernelGGL(k_interpolag<T>, grid_ij(d.nx, d.ny, d.nz), dim3(BLK), 0, stream,
                       fld[LG_U], fld[LG_V], fld[LG_W], Ftmp, Ftmp + d.field, Ftmp + 2 * d.field,
                       Ftmp + 3 * d.field, fld[LG_F_LM], fld[LG_F_MM], fld[LG_F_QN],
                       fld[LG_F_NN], g);
}

template <typename T>
void Case<T>::lagrange_sdep() {
    interpolag_sdep();
    const Geo g = geo_of(d);
    const double tf1 = 2.0, tf2 = 4.0;
    LagUpd up{};
    up.opftdelta = 1.5 * d.delta;
    up.powcoeff = -1.0 / 8.0;
    up.cnst = 2.0 * (d.delta * d.delta);
    up.tf1_2 = tf1 * tf1;
    up.tf2_2 = tf2 * tf2;
    up.ldt = d.lagran_dt;
    up.zero = 1.0e-24;
    up.expo = std::log(tf1) / (std::log(tf2) - std::log(tf1));
    up.betamin = 1.0 / (tf1 * tf2);
    const int sids[6] = {LG_S11, LG_S12, LG_S13, LG_S22, LG_S23, LG_S33};
    for (int k0 = 1; k0 <= d.nz; k0 += lag_chunk) {
        const int nc = std::min(lag_chunk, d.nz - k0 + 1);
        hipLaunchKernelGGL(k_lag_prep<T>, grid_ij(d.nx, d.ny, nc), dim3(BLK), 0, stream,
                           fld[LG_U], fld[LG_V], fld[LG_W], fld[LG_S11], fld[LG_S12],
                           fld[LG_S13], fld[LG_S22], fld[LG_S23], fld[LG_S33], lag_phys, g, k0,
                           nc);
        const long Qp = (long)nc * d.plane;
        const long Qs = (long)nc * d.splane;
        // slots 0-8 prep, 9-14 S_ij, 15-20 |S|S_ij
        const T* src[21];
        cplx<T>* dst[21];
        for (int q = 0; q < 9; ++q) src[q] = lag_phys + q * Qp;
        for (int q = 0; q < 6; ++q) src[9 + q] = fld[sids[q]] + (long)k0 * d.plane;
        for (int q = 0; q < 6; ++q) src[15 + q] = lag_phys + (9 + q) * Qp;
        for (int q = 0; q < 21; ++q) dst[q] = lag_spec + q * Qs;
        fwd2d(21, src, nullptr, d.plane, 0, dst, d.splane, 0, nc);
        const cplx<T>* isrc[42];
        int ops[42];
        T* idst[42];
        for (int q = 0; q < 21; ++q) {
            isrc[q] = dst[q];
            ops[q] = OP_G1;
            idst[q] = lag_filt + q * Qp;
            isrc[21 + q] = dst[q];
            ops[21 + q] = OP_G2;
            idst[21 + q] = lag_filt + (21 + q) * Qp;
        }
        inv2d(42, isrc, ops, d.splane, 0, idst, d.plane, 0, nc, spec, d.splane);
        hipLaunchKernelGGL(k_lag_update<T>, grid_ij(d.ld, d.ny, nc), dim3(BLK), 0, stream,
                           lag_filt, fld[LG_F_LM], fld[LG_F_MM], fld[LG_F_QN], fld[LG_F_NN],
                           fld[LG_CS_OPT2], fld[LG_BETA], fld[LG_TN_ALL], g, k0, nc, up);
    }
}

// ---------------------------------------------------------------------------
// divstress_uv.f90, divstress_w.f90
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::divstress_uv() {
    // ddx txx + ddy txy and ddx txy + ddy tyy are summed in spectral space
    // (one inverse each); the z term is added in physical space.
    const T sc = (T)(1.0 / (d.nx * d.ny));
    const int nl = d.nz - 1;
    const T* src[3] = {fld[LG_TXX], fld[LG_TXY], fld[LG_TYY]};
    const T scs[3] = {sc, sc, sc};
    cplx<T>* S[3] = {spec[0], spec[1], spec[2]};
    fwd2d(3, src, scs, d.plane, 1, S, d.splane, 1, nl);
    const cplx<T>* is[2] = {spec[0], spec[1]};
    const cplx<T>* is2[2] = {spec[1], spec[2]};
    const int ops[2] = {OP_DIVXY, OP_DIVXY};
    T* dst[2] = {tmp[0], tmp[2]};
    cplx<T>* tsp[2] = {spec[3], spec[4]};
    inv2d(2, is, ops, d.splane, 1, dst, d.plane, 1, nl, tsp, d.splane, is2);
    const Geo g = geo_of(d);
    const T rdz = (T)(1.0 / d.dz);
    hipLaunchKernelGGL(k_divt_uv<T>, grid_ij(d.ld, d.ny, nl), dim3(BLK), 0, stream, tmp[0],
                       fld[LG_TXZ], fld[LG_DIVTX], g, rdz);
    hipLaunchKernelGGL(k_divt_uv<T>, grid_ij(d.ld, d.ny, nl), dim3(BLK), 0, stream, tmp[2],
                       fld[LG_TYZ], fld[LG_DIVTY], g, rdz);
    for (int id : {LG_DIVTX, LG_DIVTY}) {
        fill_level(id, 0, BOGUS);
        fill_level(id, d.nz, BOGUS);
    }
}

template <typename T>
void Case<T>::divstress_w() {
    const T sc = (T)(1.0 / (d.nx * d.ny));
    const T* src[2] = {fld[LG_TXZ], fld[LG_TYZ]};
    const T scs[2] = {sc, sc};
    cplx<T>* S[2] = {spec[0], spec[1]};
    fwd2d(2, src, scs, d.plane, 1, S, d.splane, 1, d.nz);
    const cplx<T>* is[1] = {spec[0]};
    const cplx<T>* is2[1] = {spec[1]};
    const int ops[1] = {OP_DIVXY};
    T* dst[1] = {tmp[4]};
    cplx<T>* tsp[1] = {spec[2]};
    inv2d(1, is, ops, d.splane, 1, dst, d.plane, 1, d.nz, tsp, d.splane, is2);
    hipLaunchKernelGGL(k_divt_w<T>, grid_ij(d.ld, d.ny, d.nz), dim3(BLK), 0, stream, tmp[4],
                       fld[LG_TZZ], fld[LG_DIVTZ], geo_of(d), (T)(1.0 / d.dz));
    fill_level(LG_DIVTZ, 0, BOGUS);
}

// ---------------------------------------------------------------------------
// convec.f90
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::convec(bool reuse_spectra) {
    const T c = (T)(1.0 / (d.nx * d.ny));
    const int nl = d.nz;
    const cplx<T>* ps[6] = {spec[0], spec[1], spec[2], spec[3], spec[4], spec[5]};
    if (reuse_spectra) {
        // u, v, w spectra are the ones filt_da just produced (the filtered
        // fields' transforms); the vorticity is formed from them in spectral
        // space, with only the wall-level dudz/dvdz planes transformed.
        const T* wsrc[2] = {lev(LG_DVDZ, 1), lev(LG_DUDZ, 1)};
        const T wsc[2] = {c, c};
        cplx<T>* wd[2] = {spec[6], spec[6] + d.splane};
        fwd2d(2, wsrc, wsc, d.plane, 0, wd, d.splane, 0, 1);
        const int n = d.ny * d.NK;
        hipLaunchKernelGGL(k_vort_spec<T>, dim3((n + BLK - 1) / BLK, nl), dim3(BLK), 0, stream,
                           sU, sV, sW, spec[6], spec[6] + d.splane, spec[3], spec[4], spec[5],
                           d.ny, d.NK, d.nz, d.splane, (T)(2.0 * PI_D / d.prm.lx),
                           (T)(2.0 * PI_D / d.prm.ly), (T)(1.0 / d.dz));
        ps[0] = sU;
        ps[1] = sV;
        ps[2] = sW;
    } else {
        const T* src[3] = {fld[LG_U], fld[LG_V], fld[LG_W]};
        const T scs[3] = {c, c, c};
        cplx<T>* S[6] = {spec[0], spec[1], spec[2], spec[3], spec[4], spec[5]};
        fwd2d(3, src, scs, d.plane, 1, S, d.splane, 1, nl);
        VortArgs<T> va{fld[LG_DUDY], fld[LG_DUDZ], fld[LG_DVDX], fld[LG_DVDZ], fld[LG_DWDX],
                       fld[LG_DWDY], tmp[0], tmp[1], tmp[2], c};
        hipLaunchKernelGGL(k_vort<T>, grid_ij(d.nx, d.ny, nl), dim3(BLK), 0, stream, va,
                           geo_of(d));
        const T* vs[3] = {tmp[0], tmp[1], tmp[2]};
        fwd2d(3, vs, nullptr, d.plane, 1, S + 3, d.splane, 1, nl);
    }
    pad_inv_big(6, ps, 1, nl);
    ProdArgs<T> pa{big[0], big[1], big[2], big[3], big[4], big[5], big[6], big[7], big[8],
                   (T)(1.0 / (d.nx2 * d.ny2))};
    hipLaunchKernelGGL(k_products<T>, grid_ij(d.nx2, d.ny2, nl), dim3(BLK), 0, stream, pa,
                       d.nx2, d.ny2, d.ld_big, d.nz);
    const T* cs[3] = {big[6], big[7], big[8]};
    cplx<T>* cd[3] = {spec[0], spec[1], spec[2]};
    fwd_big_unpad(3, cs, cd, 1, nl);
    const cplx<T>* is[3] = {spec[0], spec[1], spec[2]};
    const int ops[3] = {OP_NONE, OP_NONE, OP_NONE};
    T* dst[3] = {fld[LG_RHSX], fld[LG_RHSY], fld[LG_RHSZ]};
    cplx<T>* tsp[3] = {spec[3], spec[4], spec[5]};
    inv2d(3, is, ops, d.splane, 1, dst, d.plane, 1, nl, tsp, d.splane);
    for (int id : {LG_RHSX, LG_RHSY, LG_RHSZ}) fill_level(id, 0, BOGUS);
    fill_level(LG_RHSX, d.nz, BOGUS);
    fill_level(LG_RHSY, d.nz, BOGUS);
}

// ---------------------------------------------------------------------------
// press_stag_array.f90 + tridag_array.f90
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::press_stag_array() {
    const double cst = 1.0 / (d.nx * d.ny);
    const T c2 = (T)(cst / d.tadv1 / d.prm.dt);
    {
        const T* src[2] = {fld[LG_U], fld[LG_V]};
        const T scs[2] = {c2, c2};
        cplx<T>* S[2] = {spec[0], spec[1]};
        fwd2d(2, src, scs, d.plane, 1, S, d.splane, 1, d.nz - 1);
    }
    {
        const T* src[1] = {fld[LG_W]};
        cplx<T>* S[1] = {spec[2]};
        fwd2d(1, src, &c2, d.plane, 1, S, d.splane, 1, d.nz);
    }
    {
        const T* src[2] = {lev(LG_DIVTZ, 1), lev(LG_DIVTZ, d.nz)};
        const T scs[2] = {(T)cst, (T)cst};
        cplx<T>* S[2] = {spec[3], spec[3] + d.splane};
        fwd2d(2, src, scs, d.plane, 0, S, d.splane, 0, 1);
    }
    TriGeo tg{d.ny, d.NK, d.nz, d.splane, 2.0 * PI_D / d.prm.lx, 2.0 * PI_D / d.prm.ly, d.dz,
              1.0 / (d.dz * d.dz), 1.0 / d.dz};
    const int nsys = d.ny * d.NK;
    const double tbytes = (double)d.splane * sizeof(cplx<T>) * (3.0 * d.nz + 1 + 2 + d.nz + 1);
    timer.run(stream, KK_TRIDAG, tbytes, [&] {
        hipLaunchKernelGGL(k_tridag<T>, dim3((nsys + 127) / 128), dim3(128), 0, stream, spec[0],
                           spec[1], spec[2], spec[3], spec[3] + d.splane, spec[4], gam, usc, tg,
                           dev_err);
    });
    {
        const cplx<T>* is[1] = {spec[4]};
        const int ops[1] = {OP_NONE};
        T* dst[1] = {fld[LG_P]};
        cplx<T>* tsp[1] = {spec[5]};
        inv2d(1, is, ops, d.splane, 0, dst, d.plane, 0, d.nz + 1, tsp, d.splane);
    }
    {
        const cplx<T>* is[2] = {spec[4], spec[4]};
        const int ops[2] = {OP_IKX, OP_IKY};
        T* dst[2] = {fld[LG_DPDX], fld[LG_DPDY]};
        cplx<T>* tsp[2] = {spec[5], spec[6]};
        inv2d(2, is, ops, d.splane, 1, dst, d.plane, 1, d.nz - 1, tsp, d.splane);
    }
    hipLaunchKernelGGL(k_dpdz<T>, grid_ij(d.nx, d.ny, d.nz), dim3(BLK), 0, stream, fld[LG_P],
                       fld[LG_DPDZ], geo_of(d), (T)d.dz);
    fill_level(LG_DPDX, d.nz, BOGUS);
    fill_level(LG_DPDY, d.nz, BOGUS);
}

// ---------------------------------------------------------------------------
// forcing.f90::project
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::project() {
    hipLaunchKernelGGL(k_project<T>, grid_ij(d.nx, d.ny, d.nz - 1), dim3(BLK), 0, stream,
                       fld[LG_U], fld[LG_V], fld[LG_W], fld[LG_DPDX], fld[LG_DPDY],
                       fld[LG_DPDZ], geo_of(d), (T)d.prm.dt, (T)d.tadv1);
    const unsigned nb = (unsigned)((d.plane + BLK - 1) / BLK);
    hipLaunchKernelGGL(k_copy_plane<T>, dim3(nb), dim3(BLK), 0, stream, lev(LG_U, d.nz - 1),
                       lev(LG_U, d.nz), d.plane);
    hipLaunchKernelGGL(k_copy_plane<T>, dim3(nb), dim3(BLK), 0, stream, lev(LG_V, d.nz - 1),
                       lev(LG_V, d.nz), d.plane);
    fill_level(LG_W, d.nz, 0.0);
    fill_level(LG_W, 1, 0.0);
}

// ---------------------------------------------------------------------------
// main.f90 time loop body
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::rhs_ab2() {
    const bool tb = (bool)turb;
    Ab2Args<T> a{fld[LG_U], fld[LG_V], fld[LG_W], fld[LG_RHSX], fld[LG_RHSY], fld[LG_RHSZ],
                 fld[LG_RHSX_F], fld[LG_RHSY_F], fld[LG_RHSZ_F], fld[LG_DIVTX], fld[LG_DIVTY],
                 fld[LG_DIVTZ], tb ? fld[LG_FXA] : nullptr, tb ? fld[LG_FYA] : nullptr,
                 tb ? fld[LG_FZA] : nullptr, (T)d.prm.dt, (T)d.tadv1, (T)d.tadv2,
                 (T)d.prm.mean_p_force_x, (T)d.prm.mean_p_force_y, d.prm.use_mean_p_force};
    hipLaunchKernelGGL(k_rhs_ab2<T>, grid_ij(d.ld, d.ny, d.nz), dim3(BLK), 0, stream, a,
                       geo_of(d));
    for (int id : {LG_U, LG_V, LG_W}) fill_level(id, 0, BOGUS);
    fill_level(LG_U, d.nz, BOGUS);
    fill_level(LG_V, d.nz, BOGUS);
}

template <typename T>
void Case<T>::step() {
    ++jt;
    ++jt_total;
    std::swap(fld[LG_RHSX], fld[LG_RHSX_F]);
    std::swap(fld[LG_RHSY], fld[LG_RHSY_F]);
    std::swap(fld[LG_RHSZ], fld[LG_RHSZ_F]);
    filt_da(LG_U, LG_DUDX, LG_DUDY, sU);
    filt_da(LG_V, LG_DVDX, LG_DVDY, sV);
    filt_da(LG_W, LG_DWDX, LG_DWDY, sW);
    ddz_uv(LG_U, LG_DUDZ);
    ddz_uv(LG_V, LG_DVDZ);
    ddz_w(LG_W, LG_DWDZ);
    wallstress();
    sgs_stag();
    divstress_uv();
    divstress_w();
    convec(true);
    if (turb) turbines_forcing();
    rhs_ab2();
    press_stag_array();
    rhs_gradp();
    project();
}

template <typename T>
void Case<T>::rhs_gradp() {
    hipLaunchKernelGGL(k_rhs_gradp<T>, grid_ij(d.ld, d.ny, d.nz), dim3(BLK), 0, stream,
                       fld[LG_RHSX], fld[LG_RHSY], fld[LG_RHSZ], fld[LG_DPDX], fld[LG_DPDY],
                       fld[LG_DPDZ], geo_of(d));
}

// ---------------------------------------------------------------------------
// Diagnostics
// ---------------------------------------------------------------------------
template <typename T>
double Case<T>::rmsdiv() {
    const int nb = 256;
    hipLaunchKernelGGL(k_rmsdiv_part<T>, dim3(nb), dim3(BLK), 0, stream, fld[LG_DUDX],
                       fld[LG_DVDY], fld[LG_DWDZ], geo_of(d), red);
    std::vector<double> h(nb);
    LG_HIP(hipMemcpyAsync(h.data(), red, nb * sizeof(double), hipMemcpyDeviceToHost, stream));
    LG_HIP(hipStreamSynchronize(stream));
    double s = 0.0;
    for (double x : h) s += x;
    return s / ((double)d.nx * d.ny * (d.nz - 1));
}

template <typename T>
double Case<T>::energy() {
    const int nb = 256;
    hipLaunchKernelGGL(k_energy_part<T>, dim3(nb), dim3(BLK), 0, stream, fld[LG_U], fld[LG_V],
                       fld[LG_W], geo_of(d), red);
    std::vector<double> h(nb);
    LG_HIP(hipMemcpyAsync(h.data(), red, nb * sizeof(double), hipMemcpyDeviceToHost, stream));
    LG_HIP(hipStreamSynchronize(stream));
    double s = 0.0;
    for (double x : h) s += x;
    return s * 0.5 / ((double)d.nx * d.ny * (d.nz - 1));
}

template <typename T>
double Case<T>::max_cfl() {
    const int nb = 256;
    double m[3];
    const double h3[3] = {d.dx, d.dy, d.dz};
    const int ids[3] = {LG_U, LG_V, LG_W};
    for (int q = 0; q < 3; ++q) {
        hipLaunchKernelGGL(k_maxabs_part<T>, dim3(nb), dim3(BLK), 0, stream, fld[ids[q]],
                           geo_of(d), red);
        std::vector<double> h(nb);
        LG_HIP(hipMemcpyAsync(h.data(), red, nb * sizeof(double), hipMemcpyDeviceToHost,
                              stream));
        LG_HIP(hipStreamSynchronize(stream));
        double mx = 0.0;
        for (double x : h) mx = std::max(mx, x);
        m[q] = mx / h3[q];
    }
    return d.prm.dt * std::max(m[0], std::max(m[1], m[2]));
}

template <typename T>
double Case<T>::tau_wall_bot() {
    std::vector<T> hx(d.plane), hy(d.plane);
    LG_HIP(hipMemcpyAsync(hx.data(), lev(LG_TXZ, 1), d.plane * sizeof(T), hipMemcpyDeviceToHost,
                          stream));
    LG_HIP(hipMemcpyAsync(hy.data(), lev(LG_TYZ, 1), d.plane * sizeof(T), hipMemcpyDeviceToHost,
                          stream));
    LG_HIP(hipStreamSynchronize(stream));
    double sx = 0.0, sy = 0.0;
    for (int i = 0; i < d.nx; ++i)
        for (int j = 0; j < d.ny; ++j) {
            sx += hx[(size_t)j * d.ld + i];
            sy += hy[(size_t)j * d.ld + i];
        }
    const double n = (double)d.nx * d.ny;
    return std::sqrt((sx / n) * (sx / n) + (sy / n) * (sy / n));
}

template <typename T>
void Case<T>::check_device_error() {
    int h = 0;
    LG_HIP(hipMemcpyAsync(&h, dev_err, sizeof(int), hipMemcpyDeviceToHost, stream));
    LG_HIP(hipStreamSynchronize(stream));
    if (h) {
        LG_HIP(hipMemsetAsync(dev_err, 0, sizeof(int), stream));
        set_error("tridag_array: zero pivot (tridag_array.f90:190-222)");
        throw PivotError();
    }
}

template struct Case<double>;
template struct Case<float>;

}  // namespace lg
