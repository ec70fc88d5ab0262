// Subsystem drivers: each method replaces one reference subroutine and runs
// on the case's stream over every stepped member of the batch.  See the
// kernel headers for the per-loop citations.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "case.hpp"
#include "clock.hpp"
#include "passes.hpp"
#include "lagrange.hpp"
#include "pointwise.hpp"
#include "pressure.hpp"
#include "spectral.hpp"

namespace lg {

static constexpr double PI_D = 3.1415926535897932384626433;   // param.f90:94
static constexpr int BLK = 256;

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
    if (center_lin) (void)hipFree(center_lin);
    if (center_wlin) (void)hipFree(center_wlin);
    if (center) (void)hipFree(center);
    if (fz_uv) (void)hipFree(fz_uv);
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

// Per-member step flags written in stream order from a launch's arguments
// (up to FLAGS_ARG_MAX values: 3 KB of the 4 KB argument space).
constexpr int FLAGS_ARG_MAX = 768;
struct FlagArgs {
    int n;
    int v[FLAGS_ARG_MAX];
};
// k_set_flags(int*, FlagArgs) must fit the 4 KB of kernel arguments
static_assert(sizeof(FlagArgs) + sizeof(int*) <= 4096, "k_set_flags arguments exceed 4 KB");

__global__ void k_set_flags(int* dst, FlagArgs a) {
    for (int i = threadIdx.x; i < a.n; i += blockDim.x) dst[i] = a.v[i];
}

// The Lagrangian model's fused row passes (k_lag_prep_rows, k_lag_rows):
// where nx has a kernel for them and LESGO_AMD_LAG_ROWS is not 0.  They
// need no physical quantity fields and 21 complex + 4 FP64 filtered slots
// instead of 42 real ones.
static bool lag_fused_for(int nx, int ny, int tsize) {
    const char* lr = std::getenv("LESGO_AMD_LAG_ROWS");
    if (lr && lr[0] == '0') return false;
    return tsize == 4 ? lag_rows_ok<float>(nx, ny) : lag_rows_ok<double>(nx, ny);
}

// The convective products formed in the 3/2 grid's forward row pass
// (k_rows_r2c_prod): where nx2 has a four-step split and
// LESGO_AMD_PROD_ROWS is not 0.  The three product fields are then not
// allocated.
static bool prod_fused_for(int nx2, int ny2, int tsize) {
    const char* pr = std::getenv("LESGO_AMD_PROD_ROWS");
    if (pr && pr[0] == '0') return false;
    return tsize == 4 ? prod_rows_ok<float>(nx2, ny2) : prod_rows_ok<double>(nx2, ny2);
}

// real slots of the filtered buffer: 42 fields, or 21 complex spectra (each
// within one field) and four FP64 contraction fields
static long lag_filt_slots(bool fused, int tsize) { return fused ? 21 + 4 * (8 / tsize) : 42; }

template <typename T>
Case<T>::Case(const lg_params& p, int members) : d(p), M(members), Ma(members) {
    msp = (long)(d.nz + 1) * d.splane;
    mbp = (long)(d.nz + 1) * d.bplane;
    mbsp = (long)(d.nz + 1) * d.bsplane;
    LG_HIP(hipStreamCreate(&stream));
    for (int i = 0; i < LG_NFIELDS; ++i) fld[i] = dalloc<T>(M * d.field);
    for (int i = 0; i < NSPEC; ++i) spec[i] = dalloc<cplx<T>>(M * msp);
    prod_fused = prod_fused_for(d.nx2, d.ny2, (int)sizeof(T));
    for (int i = 0; i < (prod_fused ? 6 : NBIG); ++i) big[i] = dalloc<T>(M * mbp);
    for (int i = 0; i < 6; ++i) bigspec[i] = dalloc<cplx<T>>(M * mbsp);
    for (int i = 0; i < NTMP; ++i) tmp[i] = dalloc<T>(M * d.field);
    sU = dalloc<cplx<T>>(M * msp);
    sV = dalloc<cplx<T>>(M * msp);
    sW = dalloc<cplx<T>>(M * msp);
    gam = dalloc<double>(msp);
    {
        // the elimination coefficients of the pressure systems (grid only)
        const TriGeo tg{d.ny, d.NK, d.nz, d.splane, msp, 1, 2.0 * PI_D / d.prm.lx,
                        2.0 * PI_D / d.prm.ly, d.dz, 1.0 / (d.dz * d.dz), 1.0 / d.dz};
        const int n = d.ny * d.NK;
        hipLaunchKernelGGL(k_tridag_gam, dim3((n + 255) / 256), dim3(256), 0, stream, gam, tg);
        LG_HIP(hipGetLastError());
    }
    usc = dalloc<double2>(M * msp);
    dev_err = dalloc<int>(1);
    lag_chunk = std::min(d.nz, 64);
    lag_fused = lag_fused_for(d.nx, d.ny, (int)sizeof(T));
    if (!lag_fused) lag_phys = dalloc<T>(15L * M * lag_chunk * d.plane);
    lag_spec = dalloc<cplx<T>>(21L * M * lag_chunk * d.splane);
    lag_filt = dalloc<T>(lag_filt_slots(lag_fused, (int)sizeof(T)) * M * lag_chunk * d.plane);
    Ftmp = dalloc<T>(4L * M * d.field);
    red = dalloc<double>(1024);
    d_denom = dalloc<double>(M);
    d_mpf = dalloc<double>(2L * M);
    zo.assign(M, p.zo);
    mpf.resize(2 * M);
    for (int m = 0; m < M; ++m) set_member_params(m, p.zo, p.mean_p_force_x, p.mean_p_force_y);
    mturb.resize(M);
    mjt.assign(M, 0);
    lm_init.assign(M, 0);
    hflags.assign(6 * M, 0);
    dflags_now.assign(6 * M, 0);
    d_flags = dalloc<int>(6L * M);
    mcold.assign(M, 1);
    rhsf_phys.assign(M, 1);
    cfl_mode = p.use_cfl_dt != 0;
    d_clock = dalloc<double>((long)CK_N * M);
    d_cmax = dalloc<unsigned long long>(3L * M);
    mtime.assign(M, 0.0);
    clk_fresh.assign(M, 1);
    for (int m = 0; m < M; ++m) clock_reset(m);
    // a CFL step reads the end-of-step velocity at its head, so only fixed-dt
    // cases keep the spectra between steps; LESGO_AMD_PHYSICAL_STEP=1 in the
    // environment keeps every case on the all-physical step
    const char* phys = std::getenv("LESGO_AMD_PHYSICAL_STEP");
    spec_on = !cfl_mode && !(phys && phys[0] == '1');
    if (spec_on) {
        for (int q = 0; q < 3; ++q) {
            sR[q] = dalloc<cplx<T>>(M * msp);
            sRf[q] = dalloc<cplx<T>>(M * msp);
            sF[q] = dalloc<cplx<T>>(M * msp);
            fprev[q] = dalloc<T>(M * d.field);
            tprev[q] = dalloc<T>(M * d.field);
        }
        sP = dalloc<cplx<T>>(M * msp);
        sDz = dalloc<cplx<T>>(M * msp);
    }
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
    for (int f = 0; f < 2; ++f) {
        const std::vector<T>& g = f ? g2 : g1;
        int nzc = 0, yk = 0;
        for (int j = 0; j < d.ny; ++j)
            for (int i = 0; i < d.NK; ++i)
                if (g[(size_t)j * d.NK + i] != (T)0) {
                    nzc = std::max(nzc, i + 1);
                    yk = std::max(yk, ((j < d.ny / 2) ? j : d.ny - j) + 1);
                }
        g_nzc[f] = std::max(nzc, 1);
        g_ykeep[f] = std::max(yk, 1);
    }
    G1 = dalloc<T>(d.splane);
    G2 = dalloc<T>(d.splane);
    LG_HIP(hipMemcpy(G1, g1.data(), d.splane * sizeof(T), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(G2, g2.data(), d.splane * sizeof(T), hipMemcpyHostToDevice));
}

template <typename T>
Case<T>::~Case() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    turb.reset();
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
    fr(d_denom);
    fr(d_mpf);
    fr(d_flags);
    fr(d_clock);
    fr(d_cmax);
    for (int q = 0; q < 3; ++q) {
        fr(sR[q]);
        fr(sRf[q]);
        fr(sF[q]);
        fr(fprev[q]);
        fr(tprev[q]);
    }
    fr(sP);
    fr(sDz);
    fr(tavg);
    fr(tspec);
    fr(G1);
    fr(G2);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
inline double uniform01(uint64_t seed, uint64_t fid, uint64_t i, uint64_t j, uint64_t k) {
    uint64_t key = seed * 0x100000001B3ULL;
    key = splitmix64(key ^ fid);
    key = splitmix64(key ^ ((k << 42) ^ (j << 21) ^ i));
    return (double)(key >> 11) * (1.0 / 9007199254740992.0);
}
}  // namespace

// Seeded start of member m (the vel.out-style state of initial.f90:226-239):
// log-law u with the member's z0 plus uniform noise, w = 0 at wall and lid,
// u(nz) = u(nz-1), Cs_opt2 = 0.03, everything else 0, level 0 BOGUS.
template <typename T>
void Case<T>::seed_member(int m, long seed, double noise) {
    touch(false);
    std::vector<std::vector<T>> f(4, std::vector<T>((size_t)d.field, (T)0));
    for (int k = 1; k < d.nz; ++k) {
        const double zk = (k - 0.5) * d.dz;
        const double ubar = (1.0 / d.prm.vonk) * std::log(zk / zo[m]);
        const double amp = noise * ubar;
        for (int j = 0; j < d.ny; ++j)
            for (int i = 0; i < d.nx; ++i) {
                const size_t o = ((size_t)k * d.ny + j) * d.ld + i;
                for (int q = 0; q < 3; ++q) {
                    const double r = uniform01((uint64_t)seed, q, i, j, k);
                    const double base = (q == 0) ? ubar : 0.0;
                    f[q][o] = (T)(base + amp * (2.0 * r - 1.0));
                }
            }
    }
    const size_t P = (size_t)d.plane;
    for (size_t o = 0; o < P; ++o) {
        f[2][1 * P + o] = 0;
        f[2][(size_t)d.nz * P + o] = 0;
        f[0][(size_t)d.nz * P + o] = f[0][(size_t)(d.nz - 1) * P + o];
        f[1][(size_t)d.nz * P + o] = f[1][(size_t)(d.nz - 1) * P + o];
        for (int q = 0; q < 3; ++q) f[q][o] = (T)BOGUS;
    }
    for (size_t o = P; o < (size_t)d.field; ++o) f[3][o] = (T)0.03;
    for (int id = 0; id < LG_NFIELDS; ++id)
        LG_HIP(hipMemsetAsync(mem(id, m), 0, d.field * sizeof(T), stream));
    const int ids[4] = {LG_U, LG_V, LG_W, LG_CS_OPT2};
    for (int q = 0; q < 4; ++q)
        LG_HIP(hipMemcpyAsync(mem(ids[q], m), f[q].data(), d.field * sizeof(T),
                              hipMemcpyHostToDevice, stream));
    LG_HIP(hipStreamSynchronize(stream));
    mjt[m] = 0;
    lm_init[m] = 0;
    clock_reset(m);
    mcold[m] = 1;
}

// ---------------------------------------------------------------------------
// Time-stepping clock (param: dt, tadv1, tadv2; sgs_param: lagran_dt)
// ---------------------------------------------------------------------------
// A member's clock before its first step: the run's dt with the static AB2
// weights (input_util.f90:403-409), lagran_dt = cs_count dt for a fixed dt
// and 0 for use_cfl_dt (sgs_param.f90:51), total_time 0.
template <typename T>
void Case<T>::clock_reset(int m) {
    const double dt = d.prm.dt;
    double c[CK_N] = {};
    c[CK_DT] = dt;
    c[CK_TADV1] = d.tadv1;
    c[CK_TADV2] = d.tadv2;
    c[CK_LAGDT] = cfl_mode ? 0.0 : d.lagran_dt;
    c[CK_PSCALE] = 1.0 / (d.nx * d.ny) / d.tadv1 / dt;
    LG_HIP(hipMemcpyAsync(d_clock + (long)CK_N * m, c, sizeof(c), hipMemcpyHostToDevice, stream));
    LG_HIP(hipStreamSynchronize(stream));
    mtime[m] = 0.0;
    clk_fresh[m] = 1;
}

template <typename T>
void Case<T>::clock_get(int m, double* out5) {
    if (!cfl_mode) {
        out5[0] = d.prm.dt;
        out5[1] = d.tadv1;
        out5[2] = d.tadv2;
        out5[3] = d.lagran_dt;
        out5[4] = mtime[m];
        return;
    }
    double c[CK_N];
    LG_HIP(hipMemcpyAsync(c, d_clock + (long)CK_N * m, sizeof(c), hipMemcpyDeviceToHost, stream));
    LG_HIP(hipStreamSynchronize(stream));
    for (int q = 0; q < 5; ++q) out5[q] = c[q];
}

// Restart (io.f90:62-95): total_time always; with use_cfl_dt the previous
// step's dt (the next step's dt_f) and whether the next step is Euler.
template <typename T>
void Case<T>::clock_set(int m, double dt, double total_time, bool fresh) {
    mtime[m] = total_time;
    if (!cfl_mode) return;
    double c[CK_N];
    LG_HIP(hipMemcpyAsync(c, d_clock + (long)CK_N * m, sizeof(c), hipMemcpyDeviceToHost, stream));
    LG_HIP(hipStreamSynchronize(stream));
    c[CK_DT] = dt;
    c[CK_TIME] = total_time;
    LG_HIP(hipMemcpyAsync(d_clock + (long)CK_N * m, c, sizeof(c), hipMemcpyHostToDevice, stream));
    LG_HIP(hipStreamSynchronize(stream));
    clk_fresh[m] = fresh ? 1 : 0;
}

// The use_cfl_dt head of a step: the velocity maxima of every stepped member,
// then the clock update (clock.hpp).
template <typename T>
void Case<T>::clock_begin() {
    if (!cfl_mode) return;
    const long n = (long)(d.nz - 1) * d.ny * d.nx;
    const int nb = (int)std::max(1L, std::min(1024L, n / (BLK * 32L)));
    hipLaunchKernelGGL(k_cfl_max<T>, dim3(nb, Ma), dim3(BLK), 0, stream, (const T*)fld[LG_U],
                       (const T*)fld[LG_V], (const T*)fld[LG_W], d.nx, d.ny, d.ld, d.plane,
                       d.field, d.nz, d_cmax);
    hipLaunchKernelGGL(k_clock_step, dim3(1), dim3(64), 0, stream, d_clock, d_cmax,
                       (const int*)(d_flags + 3 * M), (const int*)(d_flags + 4 * M), Ma,
                       d.prm.cfl, d.dx, d.dy, d.dz, 1.0 / (d.nx * d.ny));
}

// Bytes the constructor allocates for a batch of `members` (device arrays
// only; plan tables and filter kernels add a few kB).
size_t case_bytes(const lg_params& p, int members, int tsize) {
    const Dims d(p);
    const size_t M = (size_t)members, T = (size_t)tsize;
    const size_t msp = (size_t)(d.nz + 1) * d.splane, mbp = (size_t)(d.nz + 1) * d.bplane;
    const size_t mbsp = (size_t)(d.nz + 1) * d.bsplane;
    const size_t chunk = (size_t)std::min(d.nz, 64);
    size_t b = 0;
    b += (size_t)(LG_NFIELDS + 9 + 4) * M * d.field * T;        // fields, tmp, Ftmp
    b += (size_t)(12 + 3) * M * msp * 2 * T;                     // spec, sU/sV/sW
    const size_t nbig = prod_fused_for(d.nx2, d.ny2, tsize) ? 6 : 9;
    b += nbig * M * mbp * T + 6 * M * mbsp * 2 * T;              // big, bigspec
    b += msp * sizeof(double) + M * msp * sizeof(double2);       // gam table, usc
    const bool lf = lag_fused_for(d.nx, d.ny, tsize);
    b += ((lf ? 0 : 15) + lag_filt_slots(lf, tsize)) * M * chunk * d.plane * T +
         21 * M * chunk * d.splane * 2 * T;
    b += 3 * M * sizeof(double) + 1024 * sizeof(double) + 2 * d.splane * T;
    b += M * (CK_N * sizeof(double) + 3 * sizeof(unsigned long long) + 2 * sizeof(int));
    if (!p.use_cfl_dt) b += 11 * M * msp * 2 * T + 6 * M * d.field * T;   // spectral state
    return b;
}

template <typename T>
void Case<T>::set_member_params(int m, double zo_m, double mpfx, double mpfy) {
    zo[m] = zo_m;
    mpf[2 * m] = mpfx;
    mpf[2 * m + 1] = mpfy;
    const double den = std::log(0.5 * d.dz / zo_m);
    LG_HIP(hipMemcpyAsync(d_denom + m, &den, sizeof(double), hipMemcpyHostToDevice, stream));
    LG_HIP(hipMemcpyAsync(d_mpf + 2 * m, &mpf[2 * m], 2 * sizeof(double), hipMemcpyHostToDevice,
                          stream));
    LG_HIP(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------------------
// Transform helpers
// ---------------------------------------------------------------------------
template <typename T>
static void col_common(ColArgs<T>& c, const Dims& d, const T* G1, const T* G2) {
    c.ncols = d.NK;
    c.ny = d.ny;
    c.kxs = (T)(2.0 * PI_D / d.prm.lx);
    c.kys = (T)(2.0 * PI_D / d.prm.ly);
    c.G1 = G1;
    c.G2 = G2;
}

template <typename T>
void Case<T>::fwd2d(int njobs, const T* const* src, const T* scales, long src_plane, int src_lev0,
                    long src_ms, cplx<T>* const* dst, long dst_plane, int dst_lev0, long dst_ms,
                    int nlev, const double* mscale, int band_nzc, int band_ykeep) {
    if (nlev <= 0) return;
    launch_rows_r2c<T>(timer, px.plan, njobs, src, scales, src_plane, d.ld, src_lev0, src_ms, dst,
                       dst_plane, dst_lev0, dst_ms, d.ny, d.NK, nlev, Ma, stream, nullptr,
                       mscale, CK_N, band_nzc);
    for (int j0 = 0; j0 < njobs; j0 += MAX_JOBS) {
        const int nj = std::min(MAX_JOBS, njobs - j0);
        ColArgs<T> c{};
        c.plan = py.plan;
        for (int i = 0; i < nj; ++i) {
            c.src[i] = dst[j0 + i];
            c.dst[i] = dst[j0 + i];
            c.op[i] = OP_NONE;
            c.nzc[i] = band_nzc;
            c.ykeep[i] = band_ykeep;
        }
        c.src_plane = c.dst_plane = dst_plane;
        c.lev0_src = c.lev0_dst = dst_lev0;
        c.src_mstride = c.dst_mstride = dst_ms;
        c.map = MAP_ID;
        c.zero_nyq = 1;
        col_common(c, d, G1, G2);
        launch_cols<T, -1>(timer, c, nj, nlev, Ma, d.ny, stream, nullptr);
    }
}

template <typename T>
void Case<T>::inv2d(int njobs, const cplx<T>* const* src, const int* ops, long src_plane,
                    int src_lev0, long src_ms, T* const* dst, long dst_plane, int dst_lev0,
                    long dst_ms, int nlev, cplx<T>* const* tmpspec, const cplx<T>* const* src2) {
    if (nlev <= 0) return;
    // column-pass output goes to tmpspec[i], planes 0..nlev-1, members packed
    const long tms = (long)nlev * d.splane;
    for (int j0 = 0; j0 < njobs; j0 += MAX_JOBS) {
        const int nj = std::min(MAX_JOBS, njobs - j0);
        ColArgs<T> c{};
        c.plan = py.plan;
        // an OP_IKX job on the same input as an OP_NONE job rides along with
        // it: one y transform, two outputs (i kx is constant along y)
        // (the row pass forms i kx f from the y transform of f: the column
        // job zeroes its input's ky = ny/2 row, as the i kx job does)
        int nc = 0;
        bool taken[MAX_JOBS] = {};
        T* rdst[MAX_JOBS] = {};
        T* rdst2[MAX_JOBS] = {};
        int nce[MAX_JOBS] = {};
        for (int i = 0; i < nj; ++i) {
            if (taken[i]) continue;
            c.src[nc] = src[j0 + i];
            c.src2[nc] = src2 ? src2[j0 + i] : nullptr;
            c.dst[nc] = tmpspec[nc];
            c.op[nc] = ops[j0 + i];
            rdst[nc] = dst[j0 + i];
            const bool g1 = ops[j0 + i] == OP_G1 || ops[j0 + i] == OP_G1B;
            nce[nc] = g1 ? g_nzc[0] : ops[j0 + i] == OP_G2 ? g_nzc[1] : 0;
            if (g1 || ops[j0 + i] == OP_G2) {
                // test-filtered: nothing outside the filter's band
                const int f = ops[j0 + i] == OP_G2 ? 1 : 0;
                c.nzc[nc] = g_nzc[f];
                c.ykeep[nc] = g_ykeep[f];
            }
            if (ops[j0 + i] == OP_NONE && !c.src2[nc]) {
                for (int i2 = i + 1; i2 < nj; ++i2) {
                    if (!taken[i2] && ops[j0 + i2] == OP_IKX && src[j0 + i2] == src[j0 + i] &&
                        !(src2 && src2[j0 + i2])) {
                        c.nyq0[nc] = 1;
                        rdst2[nc] = dst[j0 + i2];
                        taken[i2] = true;
                        break;
                    }
                }
            }
            ++nc;
        }
        c.src_plane = src_plane;
        c.dst_plane = d.splane;
        c.lev0_src = src_lev0;
        c.lev0_dst = 0;
        c.src_mstride = src_ms;
        c.dst_mstride = tms;
        c.map = MAP_ID;
        c.zero_nyq = 0;
        col_common(c, d, G1, G2);
        launch_cols<T, +1>(timer, c, nc, nlev, Ma, d.ny, stream, nullptr);
        std::vector<const cplx<T>*> ts(tmpspec, tmpspec + nc);
        launch_rows_c2r<T>(timer, px.plan, nc, ts.data(), d.splane, 0, tms, rdst, dst_plane,
                           d.ld, dst_lev0, dst_ms, d.ny, d.NK, nlev, Ma, stream, nullptr, nce,
                           rdst2, (T)(2.0 * PI_D / d.prm.lx));
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
    c.src_mstride = msp;
    c.dst_mstride = mbsp;
    c.map = MAP_PAD;
    c.zero_nyq = 0;
    col_common(c, d, G1, G2);
    launch_cols<T, +1>(timer, c, njobs, nlev, Ma, d.ny, stream, nullptr);
    std::vector<const cplx<T>*> bs(bigspec, bigspec + njobs);
    launch_rows_c2r<T>(timer, pbx.plan, njobs, bs.data(), d.bsplane, lev0, mbsp, big, d.bplane,
                       d.ld_big, lev0, mbp, d.ny2, d.NK, nlev, Ma, stream, nullptr);
}

template <typename T>
void Case<T>::fwd_big_unpad(int njobs, const T* const* src, cplx<T>* const* dst, int lev0,
                            int nlev) {
    launch_rows_r2c<T>(timer, pbx.plan, njobs, src, nullptr, d.bplane, d.ld_big, lev0, mbp,
                       bigspec, d.bsplane, lev0, mbsp, d.ny2, d.NK, nlev, Ma, stream, nullptr);
    cols_unpad(njobs, dst, lev0, nlev);
}

// The forward column pass of fwd_big_unpad: bigspec[0..njobs) into dst.
template <typename T>
void Case<T>::cols_unpad(int njobs, cplx<T>* const* dst, int lev0, int nlev) {
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
    c.src_mstride = mbsp;
    c.dst_mstride = msp;
    c.map = MAP_UNPAD;
    c.zero_nyq = 1;
    col_common(c, d, G1, G2);
    launch_cols<T, -1>(timer, c, njobs, nlev, Ma, d.ny2, stream, nullptr);
}

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
#define GEO(k0, nlev) (Geo{d.nx, d.ny, d.nz, d.ld, d.plane, (k0), (nlev), d.field})

static inline dim3 grid_ij(int W, int ny, int nlev, int M) {
    return dim3((unsigned)((W * ny + BLK - 1) / BLK), (unsigned)(nlev * M));
}

template <typename T>
void Case<T>::fill_level(int id, int k, double v) {
    T* p = fld[id] + (long)k * d.plane;
    hipLaunchKernelGGL(k_fill_planes<T>, dim3((unsigned)((d.plane + BLK - 1) / BLK), Ma),
                       dim3(BLK), 0, stream, p, d.field, d.plane, (T)v);
}

template <typename T>
void Case<T>::fill_levels(std::initializer_list<Fill> f) {
    FillList<T> a{};
    int n = 0;
    for (const Fill& x : f) {
        a.p[n] = fld[x.id] + (long)x.k * d.plane;
        a.v[n] = (T)x.v;
        if (++n == MAX_FILLS) break;
    }
    if (n == 0) return;
    a.mstride = d.field;
    a.plane = d.plane;
    hipLaunchKernelGGL(k_fill_list<T>, dim3((unsigned)((d.plane + BLK - 1) / BLK), Ma, n),
                       dim3(BLK), 0, stream, a);
}

// dst[q] = src[q], q < n (a kernel, so that captured steps hold no copy nodes)
template <typename T>
static void copy_dev(const T* src, T* dst, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_plane<T>, dim3((unsigned)((n + BLK - 1) / BLK), 1), dim3(BLK), 0, s,
                       src, dst, 0L, n);
}

template <typename T>
void Case<T>::copy_member(int src, int dst) {
    if (src == dst) return;
    touch(false);
    for (int id = 0; id < LG_NFIELDS; ++id)
        LG_HIP(hipMemcpyAsync(mem(id, dst), mem(id, src), d.field * sizeof(T),
                              hipMemcpyDeviceToDevice, stream));
    set_member_params(dst, zo[src], mpf[2 * src], mpf[2 * src + 1]);
    mjt[dst] = mjt[src];
    lm_init[dst] = lm_init[src];
    LG_HIP(hipMemcpyAsync(d_clock + (long)CK_N * dst, d_clock + (long)CK_N * src,
                          CK_N * sizeof(double), hipMemcpyDeviceToDevice, stream));
    mtime[dst] = mtime[src];
    clk_fresh[dst] = clk_fresh[src];
    copy_spec_member(src, dst);
    if (turb || mturb[dst].n() || mturb[src].n()) {
        pull_turbine_state();
        mturb[dst] = mturb[src];
        rebuild_turbines();
    }
}

template <typename T>
void Case<T>::retire_member(int slot) {
    touch(false);
    const int last = Ma - 1;
    const bool disks = mturb[slot].n() > 0 || mturb[last].n() > 0;
    if (disks) pull_turbine_state();
    if (slot != last) {
        for (int id = 0; id < LG_NFIELDS; ++id)
            LG_HIP(hipMemcpyAsync(mem(id, slot), mem(id, last), d.field * sizeof(T),
                                  hipMemcpyDeviceToDevice, stream));
        set_member_params(slot, zo[last], mpf[2 * last], mpf[2 * last + 1]);
        mturb[slot] = mturb[last];
        mjt[slot] = mjt[last];
        lm_init[slot] = lm_init[last];
        LG_HIP(hipMemcpyAsync(d_clock + (long)CK_N * slot, d_clock + (long)CK_N * last,
                              CK_N * sizeof(double), hipMemcpyDeviceToDevice, stream));
        mtime[slot] = mtime[last];
        clk_fresh[slot] = clk_fresh[last];
        copy_spec_member(last, slot);
    }
    mturb[last] = MemberTurbines{};
    Ma = last;
    if (disks) rebuild_turbines();
    LG_HIP(hipStreamSynchronize(stream));
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
    fwd2d(1, src, &sc, d.plane, 0, d.field, S0, d.splane, 0, msp, d.nz + 1);
    const cplx<T>* s3[3] = {sp, sp, sp};
    const int ops[3] = {OP_NONE, OP_IKX, OP_IKY};
    T* dst[3] = {fld[f], fld[dfdx], fld[dfdy]};
    cplx<T>* tsp[3] = {spec[1], spec[2], spec[3]};
    inv2d(3, s3, ops, d.splane, 0, msp, dst, d.plane, 0, d.field, d.nz + 1, tsp);
}

// filt_da (derivatives.f90:166-211) of u, v and w: the three forward
// transforms in one pass pair, the nine inverses (filtered field, d/dx, d/dy
// of each) in another; the spectra are kept in sU, sV, sW for convec.
template <typename T>
void Case<T>::filt_da_uvw(bool forward) {
    const T sc = (T)(1.0 / (d.nx * d.ny));
    const T* src[3] = {fld[LG_U], fld[LG_V], fld[LG_W]};
    const T scs[3] = {sc, sc, sc};
    cplx<T>* S[3] = {sU, sV, sW};
    if (forward) fwd2d(3, src, scs, d.plane, 0, d.field, S, d.splane, 0, msp, d.nz + 1);
    const cplx<T>* is[9] = {sU, sU, sU, sV, sV, sV, sW, sW, sW};
    const int ops[9] = {OP_NONE, OP_IKX, OP_IKY, OP_NONE, OP_IKX, OP_IKY,
                        OP_NONE, OP_IKX, OP_IKY};
    T* dst[9] = {fld[LG_U], fld[LG_DUDX], fld[LG_DUDY], fld[LG_V], fld[LG_DVDX],
                 fld[LG_DVDY], fld[LG_W], fld[LG_DWDX], fld[LG_DWDY]};
    inv2d(9, is, ops, d.splane, 0, msp, dst, d.plane, 0, d.field, d.nz + 1, spec);
}

template <typename T>
void Case<T>::ddxy(int f, int dfdx, int dfdy) {
    const T sc = (T)(1.0 / (d.nx * d.ny));
    const T* src[1] = {fld[f]};
    cplx<T>* S0[1] = {spec[0]};
    fwd2d(1, src, &sc, d.plane, 0, d.field, S0, d.splane, 0, msp, d.nz + 1);
    const cplx<T>* s2[2] = {spec[0], spec[0]};
    int ops[2] = {OP_IKX, OP_IKY};
    T* dst[2] = {dfdx >= 0 ? fld[dfdx] : nullptr, dfdy >= 0 ? fld[dfdy] : nullptr};
    cplx<T>* tsp[2] = {spec[1], spec[2]};
    int nj = 2;
    if (dfdx < 0) {
        ops[0] = OP_IKY;
        dst[0] = fld[dfdy];
        nj = 1;
    } else if (dfdy < 0) {
        nj = 1;
    }
    inv2d(nj, s2, ops, d.splane, 0, msp, dst, d.plane, 0, d.field, d.nz + 1, tsp);
}

template <typename T>
void Case<T>::ddx(int f, int dfdx) { ddxy(f, dfdx, -1); }

template <typename T>
void Case<T>::ddy(int f, int dfdy) { ddxy(f, -1, dfdy); }

template <typename T>
void Case<T>::ddz_uv(int f, int dfdz) {
    hipLaunchKernelGGL(k_ddz_uv<T>, grid_ij(d.nx, d.ny, d.nz, Ma), dim3(BLK), 0, stream, fld[f],
                       fld[dfdz], GEO(1, d.nz), (T)(1.0 / d.dz));
    fill_levels({{dfdz, 0, BOGUS}, {dfdz, 1, BOGUS}, {dfdz, d.nz, BOGUS}});
}

template <typename T>
void Case<T>::ddz_w(int f, int dfdz) {
    hipLaunchKernelGGL(k_ddz_w<T>, grid_ij(d.nx, d.ny, d.nz, Ma), dim3(BLK), 0, stream, fld[f],
                       fld[dfdz], GEO(0, d.nz), (T)(1.0 / d.dz));
    fill_levels({{dfdz, 0, BOGUS}, {dfdz, d.nz, BOGUS}});
}

template <typename T>
void Case<T>::ddz_uvw() {
    if (d.nz >= 2) {
        const int nch = (d.nz - 1 + DDZ_L - 1) / DDZ_L;
        hipLaunchKernelGGL(k_ddz_uvw<T>, grid_ij((d.nx + 1) / 2, d.ny, nch, Ma), dim3(BLK), 0, stream,
                           (const T*)fld[LG_U], (const T*)fld[LG_V], (const T*)fld[LG_W],
                           fld[LG_DUDZ], fld[LG_DVDZ], fld[LG_DWDZ], GEO(1, nch),
                           (T)(1.0 / d.dz));
    }
    fill_levels({{LG_DUDZ, 0, BOGUS}, {LG_DUDZ, 1, BOGUS}, {LG_DUDZ, d.nz, BOGUS},
                 {LG_DVDZ, 0, BOGUS}, {LG_DVDZ, 1, BOGUS}, {LG_DVDZ, d.nz, BOGUS},
                 {LG_DWDZ, 0, BOGUS}, {LG_DWDZ, d.nz, BOGUS}});
}

// One plane per member (member stride = field) filtered by op.
template <typename T>
void Case<T>::test_filter_plane(const T* in, T* out, int op) {
    const T* src[1] = {in};
    const T one = (T)1;
    cplx<T>* S0[1] = {spec[4]};
    fwd2d(1, src, &one, d.plane, 0, d.field, S0, d.splane, 0, msp, 1);
    const cplx<T>* s1[1] = {spec[4]};
    T* dst[1] = {out};
    cplx<T>* tsp[1] = {spec[5]};
    inv2d(1, s1, &op, d.splane, 0, msp, dst, d.plane, 0, d.field, 1, tsp);
}

// ---------------------------------------------------------------------------
// wallstress.f90 (lbc_mom = 2, ubc_mom = 0)
// ---------------------------------------------------------------------------
// su, sv (the spectral-resident step): the velocity spectra, normalised by
// 1/(nx ny), whose inverse is the step's u, v; the test filter of the wall
// level is then their G_test band, with no forward transform (the same
// values to rounding: G_test is 1/(nx ny) inside a sharp band and zero
// outside it and on the Nyquist row; the band mask also drops the kx Nyquist
// column, which the physical u, v do not carry).
template <typename T>
void Case<T>::wallstress(const cplx<T>* su, const cplx<T>* sv) {
    const T one = (T)1;
    const cplx<T>* s2[2] = {spec[4], spec[5]};
    int ops[2] = {OP_G1, OP_G1};
    if (su && sv) {
        s2[0] = su + d.splane;
        s2[1] = sv + d.splane;
        ops[0] = ops[1] = OP_G1B;
    } else {
        const T* src[2] = {lev(LG_U, 1), lev(LG_V, 1)};
        const T sc[2] = {one, one};
        cplx<T>* S0[2] = {spec[4], spec[5]};
        fwd2d(2, src, sc, d.plane, 0, d.field, S0, d.splane, 0, msp, 1);
    }
    T* dst[2] = {tmp[0], tmp[1]};
    cplx<T>* tsp[2] = {spec[6], spec[7]};
    inv2d(2, s2, ops, d.splane, 0, msp, dst, d.plane, 0, d.field, 1, tsp);
    hipLaunchKernelGGL(k_wall_lbc<T>, grid_ij(d.nx, d.ny, 1, Ma), dim3(BLK), 0, stream,
                       (const T*)tmp[0], (const T*)tmp[1], d.field, (const T*)fld[LG_U],
                       (const T*)fld[LG_V], fld[LG_TXZ], fld[LG_TYZ], fld[LG_DUDZ], fld[LG_DVDZ],
                       GEO(1, 1), (const double*)d_denom, d.prm.vonk, d.dz);
    fill_levels({{LG_TXZ, d.nz, 0.0}, {LG_TYZ, d.nz, 0.0}, {LG_DUDZ, d.nz, 0.0},
                 {LG_DVDZ, d.nz, 0.0}});
}

// ---------------------------------------------------------------------------
// sgs_stag_util.f90
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::calc_sij() {
    SijArgs<T> a{fld[LG_DUDX], fld[LG_DUDY], fld[LG_DUDZ], fld[LG_DVDX], fld[LG_DVDY],
                 fld[LG_DVDZ], fld[LG_DWDX], fld[LG_DWDY], fld[LG_DWDZ], fld[LG_S11],
                 fld[LG_S12], fld[LG_S13], fld[LG_S22], fld[LG_S23], fld[LG_S33]};
    hipLaunchKernelGGL(k_calc_sij<T>, grid_ij(d.nx, d.ny, d.nz, Ma), dim3(BLK), 0, stream, a,
                       GEO(1, d.nz));
}

template <typename T>
void Case<T>::sgs_stag() {
    // without a dynamic-model update between them, S_ij and the stresses
    // are formed in one pass (k_sij_tau); Cs_opt2 is set first either way
    if (lag_now) calc_sij();
    // Cs_opt2 = 0.03 at a fresh member's first step (sgs_stag_util.f90:187-189)
    bool any_cs = false;
    for (int m = 0; m < Ma; ++m) any_cs |= hflags[M + m] != 0;
    if (any_cs)
        hipLaunchKernelGGL(k_fill_members_if<T>,
                           dim3((unsigned)((d.field + BLK - 1) / BLK), Ma), dim3(BLK), 0, stream,
                           fld[LG_CS_OPT2], (const int*)(d_flags + M), d.field, d.field,
                           (T)0.03);
    if (lag_now) lagrange_sdep();
    TauArgs<T> a{fld[LG_S11], fld[LG_S12], fld[LG_S13], fld[LG_S22], fld[LG_S23], fld[LG_S33],
                 fld[LG_CS_OPT2], fld[LG_NU_T], fld[LG_TXX], fld[LG_TXY], fld[LG_TYY],
                 fld[LG_TXZ], fld[LG_TYZ], fld[LG_TZZ], (T)(d.delta * d.delta), (T)d.prm.nu};
    if (lag_now) {
        hipLaunchKernelGGL(k_sgs_tau<T>, grid_ij(d.nx, d.ny, d.nz, Ma), dim3(BLK), 0, stream, a,
                           GEO(1, d.nz));
    } else {
        SijArgs<T> sa{fld[LG_DUDX], fld[LG_DUDY], fld[LG_DUDZ], fld[LG_DVDX], fld[LG_DVDY],
                      fld[LG_DVDZ], fld[LG_DWDX], fld[LG_DWDY], fld[LG_DWDZ], fld[LG_S11],
                      fld[LG_S12], fld[LG_S13], fld[LG_S22], fld[LG_S23], fld[LG_S33]};
        // a spectral-resident case leaves S_ij and Nu_t to be made on demand
        const int nch = (d.nz + SIJ_L - 1) / SIJ_L;
        hipLaunchKernelGGL(k_sij_tau<T>, grid_ij(d.nx, d.ny, nch, Ma), dim3(BLK), 0, stream,
                           sa, a, GEO(1, d.nz), spec_on ? 0 : 1);
    }
    if (spec_on) {
        // (after_body does the same for whole steps, replayed graphs included)
        if (lag_now) stale &= ~(unsigned)SG_SIJ;
        else stale |= SG_SIJ;
    }
    fill_levels({{LG_TXX, 0, BOGUS}, {LG_TXY, 0, BOGUS}, {LG_TXZ, 0, BOGUS}, {LG_TYY, 0, BOGUS},
                 {LG_TYZ, 0, BOGUS}, {LG_TZZ, 0, BOGUS}, {LG_TXX, d.nz, BOGUS},
                 {LG_TXY, d.nz, BOGUS}, {LG_TYY, d.nz, BOGUS}, {LG_TZZ, d.nz, BOGUS}});
}

// ---------------------------------------------------------------------------
// interpolag_Sdep.f90 / lagrange_Sdep.f90
// ---------------------------------------------------------------------------
// in_place: F_* hold the interpolated values afterwards (the reference's
// interpolag_Sdep; the F_* are copied to Ftmp and gathered back).  Else the
// interpolated values go to Ftmp (F_* unchanged), where lagrange_Sdep's
// update reads them: no copies.
template <typename T>
void Case<T>::interpolag_sdep(bool in_place) {
    const int ids[4] = {LG_F_LM, LG_F_MM, LG_F_QN, LG_F_NN};
    const long Fq = (long)M * d.field;
    const T* src[4];
    T* dst[4];
    for (int q = 0; q < 4; ++q) {
        if (in_place) {
            copy_dev<T>(fld[ids[q]], Ftmp + q * Fq, Ma * d.field, stream);
            src[q] = Ftmp + q * Fq;
            dst[q] = fld[ids[q]];
        } else {
            src[q] = fld[ids[q]];
            dst[q] = Ftmp + q * Fq;
        }
    }
    InterpGeo g{d.nx, d.ny, d.nz, d.ld, d.plane, d.dx, d.dy, d.dz,
                d.prm.lx, d.prm.ly, d.prm.lz, d.lagran_dt, d.field, clk()};
    hipLaunchKernelGGL(k_interpolag<T>, grid_ij(d.nx, d.ny, d.nz, Ma), dim3(BLK), 0, stream,
                       fld[LG_U], fld[LG_V], fld[LG_W], src[0], src[1], src[2], src[3], dst[0],
                       dst[1], dst[2], dst[3], g, (const int*)(d_flags + 2 * M));
}

template <typename T>
void Case<T>::lagrange_sdep() {
    interpolag_sdep(false);
    const Geo g = GEO(1, d.nz);
    const double tf1 = 2.0, tf2 = 4.0;
    LagUpd up{};
    up.opftdelta = 1.5 * d.delta;
    up.powcoeff = -1.0 / 8.0;
    up.cnst = 2.0 * (d.delta * d.delta);
    up.tf1_2 = tf1 * tf1;
    up.tf2_2 = tf2 * tf2;
    up.ldt = d.lagran_dt;
    up.clk = clk();
    up.zero = 1.0e-24;
    up.expo = std::log(tf1) / (std::log(tf2) - std::log(tf1));
    up.betamin = 1.0 / (tf1 * tf2);
    const int sids[6] = {LG_S11, LG_S12, LG_S13, LG_S22, LG_S23, LG_S33};
    for (int k0 = 1; k0 <= d.nz; k0 += lag_chunk) {
        const int nc = std::min(lag_chunk, d.nz - k0 + 1);
        // slot q of a chunk buffer: Ma members of nc planes each
        const long Qp = (long)Ma * nc * d.plane, mp = (long)nc * d.plane;
        const long Qs = (long)Ma * nc * d.splane, ms = (long)nc * d.splane;
        // spectral slots 0-8 prep, 9-14 S_ij, 15-20 |S| S_ij
        cplx<T>* dst[21];
        for (int q = 0; q < 21; ++q) dst[q] = lag_spec + q * Qs;
        if (lag_fused) {
            // the 21 quantities formed and transformed along x in one pass
            // (k_lag_prep_rows), then the column pass over the G_test band
            LagPrepArgs<T> pa{};
            pa.plan = px.plan;
            pa.u = fld[LG_U];
            pa.v = fld[LG_V];
            pa.w = fld[LG_W];
            pa.S11 = fld[LG_S11];
            pa.S12 = fld[LG_S12];
            pa.S13 = fld[LG_S13];
            pa.S22 = fld[LG_S22];
            pa.S23 = fld[LG_S23];
            pa.S33 = fld[LG_S33];
            pa.fplane = d.plane;
            pa.fmstride = d.field;
            pa.ld = d.ld;
            pa.nz = d.nz;
            pa.k0 = k0;
            pa.nlev = nc;
            pa.out = lag_spec;
            pa.qstride = Qs;
            pa.plane = d.splane;
            pa.mstride = ms;
            pa.ncols = d.NK;
            pa.nwr = g_nzc[0];
            if (!launch_lag_prep_rows<T>(timer, pa, Ma, stream))
                throw std::runtime_error("k_lag_prep_rows not launchable for this grid");
            for (int q0 = 0; q0 < 21; q0 += MAX_JOBS) {
                const int nj = std::min(MAX_JOBS, 21 - q0);
                ColArgs<T> cc{};
                cc.plan = py.plan;
                for (int i = 0; i < nj; ++i) {
                    cc.src[i] = cc.dst[i] = dst[q0 + i];
                    cc.op[i] = OP_NONE;
                    cc.nzc[i] = g_nzc[0];
                    cc.ykeep[i] = g_ykeep[0];
                }
                cc.src_plane = cc.dst_plane = d.splane;
                cc.lev0_src = cc.lev0_dst = 0;
                cc.src_mstride = cc.dst_mstride = ms;
                cc.map = MAP_ID;
                cc.zero_nyq = 1;
                col_common(cc, d, G1, G2);
                launch_cols<T, -1>(timer, cc, nj, nc, Ma, d.ny, stream, nullptr);
            }
        } else {
            hipLaunchKernelGGL(k_lag_prep<T>, grid_ij(d.nx, d.ny, nc, Ma), dim3(BLK), 0, stream,
                               fld[LG_U], fld[LG_V], fld[LG_W], fld[LG_S11], fld[LG_S12],
                               fld[LG_S13], fld[LG_S22], fld[LG_S23], fld[LG_S33], lag_phys, g,
                               k0, nc, Ma);
            const T* psrc[15];
            cplx<T>* pdst[15];
            for (int q = 0; q < 9; ++q) {
                psrc[q] = lag_phys + q * Qp;
                pdst[q] = dst[q];
            }
            for (int q = 0; q < 6; ++q) {
                psrc[9 + q] = lag_phys + (9 + q) * Qp;
                pdst[9 + q] = dst[15 + q];
            }
            // only the test filters read these spectra: the G1 band (G2's lies
            // inside it)
            fwd2d(15, psrc, nullptr, d.plane, 0, mp, pdst, d.splane, 0, ms, nc, nullptr, g_nzc[0],
                  g_ykeep[0]);
            const T* ssrc[6];
            for (int q = 0; q < 6; ++q) ssrc[q] = fld[sids[q]];
            fwd2d(6, ssrc, nullptr, d.plane, k0, d.field, dst + 9, d.splane, 0, ms, nc, nullptr,
                  g_nzc[0], g_ykeep[0]);
        }
        double* contr = nullptr;
        if (lag_fused) {
            // per filter: the 21 quantities' column pass into lag_filt
            // (quantity q at q * Qs complex), then k_lag_rows forms L_ij M_ij
            // and M_ij M_ij (Q_ij N_ij and N_ij N_ij) in four FP64 fields Qp
            // apart after them
            cplx<T>* colb = reinterpret_cast<cplx<T>*>(lag_filt);
            contr = reinterpret_cast<double*>(lag_filt + 21 * Qp);
            for (int f = 0; f < 2; ++f) {
                for (int q0 = 0; q0 < 21; q0 += MAX_JOBS) {
                    const int nj = std::min(MAX_JOBS, 21 - q0);
                    ColArgs<T> cc{};
                    cc.plan = py.plan;
                    for (int i = 0; i < nj; ++i) {
                        cc.src[i] = dst[q0 + i];
                        cc.dst[i] = colb + (q0 + i) * Qs;
                        cc.op[i] = f ? OP_G2 : OP_G1;
                        cc.nzc[i] = g_nzc[f];
                        cc.ykeep[i] = g_ykeep[f];
                    }
                    cc.src_plane = cc.dst_plane = d.splane;
                    cc.lev0_src = cc.lev0_dst = 0;
                    cc.src_mstride = cc.dst_mstride = ms;
                    cc.map = MAP_ID;
                    cc.zero_nyq = 0;
                    col_common(cc, d, G1, G2);
                    launch_cols<T, +1>(timer, cc, nj, nc, Ma, d.ny, stream, nullptr);
                }
                LagRowArgs<T> la{};
                la.plan = px.plan;
                la.src = colb;
                la.qstride = Qs;
                la.plane = d.splane;
                la.mstride = ms;
                la.nrows = d.ny;
                la.ncols = d.NK;
                la.nce = g_nzc[f];
                la.nlev = nc;
                la.out0 = contr + (2 * f) * Qp;
                la.out1 = contr + (2 * f + 1) * Qp;
                la.oplane = d.plane;
                la.omstride = mp;
                la.ld = d.ld;
                la.cn = up.cnst;
                la.tf2 = f ? up.tf2_2 : up.tf1_2;
                if (!launch_lag_rows<T>(timer, la, Ma, stream))
                    throw std::runtime_error("k_lag_rows not launchable for this grid");
            }
        } else {
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
            inv2d(42, isrc, ops, d.splane, 0, ms, idst, d.plane, 0, mp, nc, spec);
        }
        hipLaunchKernelGGL(k_lag_update<T>, grid_ij(d.ld, d.ny, nc, Ma), dim3(BLK), 0, stream,
                           (const T*)lag_filt, fld[LG_F_LM], fld[LG_F_MM], fld[LG_F_QN],
                           fld[LG_F_NN], fld[LG_CS_OPT2], fld[LG_BETA], fld[LG_TN_ALL], g, k0, nc,
                           Ma, up, (const int*)(d_flags + 2 * M), (const double*)contr,
                           (const T*)Ftmp, (long)M * d.field);
    }
    if (cfl_mode)
        hipLaunchKernelGGL(k_clock_lag_reset, dim3(1), dim3(64), 0, stream, d_clock,
                           (const int*)(d_flags + 2 * M), Ma);
    for (int m = 0; m < Ma; ++m)
        if (hflags[2 * M + m] == 2) lm_init[m] = 1;
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
    fwd2d(3, src, scs, d.plane, 1, d.field, S, d.splane, 1, msp, nl);
    const cplx<T>* is[2] = {spec[0], spec[1]};
    const cplx<T>* is2[2] = {spec[1], spec[2]};
    const int ops[2] = {OP_DIVXY, OP_DIVXY};
    T* dst[2] = {tmp[0], tmp[2]};
    cplx<T>* tsp[2] = {spec[3], spec[4]};
    inv2d(2, is, ops, d.splane, 1, msp, dst, d.plane, 1, d.field, nl, tsp, is2);
    const T rdz = (T)(1.0 / d.dz);
    hipLaunchKernelGGL(k_divt_uv<T>, grid_ij(d.ld, d.ny, nl, Ma), dim3(BLK), 0, stream,
                       (const T*)tmp[0], (const T*)fld[LG_TXZ], fld[LG_DIVTX], GEO(1, nl), rdz);
    hipLaunchKernelGGL(k_divt_uv<T>, grid_ij(d.ld, d.ny, nl, Ma), dim3(BLK), 0, stream,
                       (const T*)tmp[2], (const T*)fld[LG_TYZ], fld[LG_DIVTY], GEO(1, nl), rdz);
    fill_levels({{LG_DIVTX, 0, BOGUS}, {LG_DIVTX, d.nz, BOGUS}, {LG_DIVTY, 0, BOGUS},
                 {LG_DIVTY, d.nz, BOGUS}});
}

template <typename T>
void Case<T>::divstress_w() {
    const T sc = (T)(1.0 / (d.nx * d.ny));
    const T* src[2] = {fld[LG_TXZ], fld[LG_TYZ]};
    const T scs[2] = {sc, sc};
    cplx<T>* S[2] = {spec[0], spec[1]};
    fwd2d(2, src, scs, d.plane, 1, d.field, S, d.splane, 1, msp, d.nz);
    const cplx<T>* is[1] = {spec[0]};
    const cplx<T>* is2[1] = {spec[1]};
    const int ops[1] = {OP_DIVXY};
    T* dst[1] = {tmp[4]};
    cplx<T>* tsp[1] = {spec[2]};
    inv2d(1, is, ops, d.splane, 1, msp, dst, d.plane, 1, d.field, d.nz, tsp, is2);
    hipLaunchKernelGGL(k_divt_w<T>, grid_ij(d.ld, d.ny, d.nz, Ma), dim3(BLK), 0, stream,
                       (const T*)tmp[4], (const T*)fld[LG_TZZ], fld[LG_DIVTZ], GEO(1, d.nz),
                       (T)(1.0 / d.dz));
    fill_levels({{LG_DIVTZ, 0, BOGUS}});
}

// divstress_uv + divstress_w with the transforms of both in one forward and
// one inverse pass over levels 1..nz (the uv sums' level nz is computed and
// not used).
template <typename T>
void Case<T>::divstress_uvw() {
    const T sc = (T)(1.0 / (d.nx * d.ny));
    const T* src[5] = {fld[LG_TXX], fld[LG_TXY], fld[LG_TYY], fld[LG_TXZ], fld[LG_TYZ]};
    const T scs[5] = {sc, sc, sc, sc, sc};
    cplx<T>* S[5] = {spec[0], spec[1], spec[2], spec[3], spec[4]};
    fwd2d(5, src, scs, d.plane, 1, d.field, S, d.splane, 1, msp, d.nz);
    const cplx<T>* is[3] = {spec[0], spec[1], spec[3]};
    const cplx<T>* is2[3] = {spec[1], spec[2], spec[4]};
    const int ops[3] = {OP_DIVXY, OP_DIVXY, OP_DIVXY};
    T* dst[3] = {tmp[0], tmp[2], tmp[4]};
    cplx<T>* tsp[3] = {spec[5], spec[6], spec[7]};
    inv2d(3, is, ops, d.splane, 1, msp, dst, d.plane, 1, d.field, d.nz, tsp, is2);
    const int nl = d.nz - 1;
    const T rdz = (T)(1.0 / d.dz);
    hipLaunchKernelGGL(k_divt_uv<T>, grid_ij(d.ld, d.ny, nl, Ma), dim3(BLK), 0, stream,
                       (const T*)tmp[0], (const T*)fld[LG_TXZ], fld[LG_DIVTX], GEO(1, nl), rdz);
    hipLaunchKernelGGL(k_divt_uv<T>, grid_ij(d.ld, d.ny, nl, Ma), dim3(BLK), 0, stream,
                       (const T*)tmp[2], (const T*)fld[LG_TYZ], fld[LG_DIVTY], GEO(1, nl), rdz);
    hipLaunchKernelGGL(k_divt_w<T>, grid_ij(d.ld, d.ny, d.nz, Ma), dim3(BLK), 0, stream,
                       (const T*)tmp[4], (const T*)fld[LG_TZZ], fld[LG_DIVTZ], GEO(1, d.nz), rdz);
    fill_levels({{LG_DIVTX, 0, BOGUS}, {LG_DIVTX, d.nz, BOGUS}, {LG_DIVTY, 0, BOGUS},
                 {LG_DIVTY, d.nz, BOGUS}, {LG_DIVTZ, 0, BOGUS}});
}

// ---------------------------------------------------------------------------
// convec.f90
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::convec(bool reuse_spectra, bool spectral_out) {
    const T c = (T)(1.0 / (d.nx * d.ny));
    const int nl = d.nz;
    const cplx<T>* ps[6] = {spec[0], spec[1], spec[2], spec[3], spec[4], spec[5]};
    if (reuse_spectra) {
        // u, v, w spectra are the ones filt_da just produced (the filtered
        // fields' transforms); the vorticity is formed from them in spectral
        // space, with only the wall-level dvdz/dudz planes transformed (into
        // planes 0 and 1 of spec[6]).
        const T* wsrc[2] = {lev(LG_DVDZ, 1), lev(LG_DUDZ, 1)};
        const T wsc[2] = {c, c};
        cplx<T>* wd[2] = {spec[6], spec[6] + d.splane};
        fwd2d(2, wsrc, wsc, d.plane, 0, d.field, wd, d.splane, 0, msp, 1);
        const int n = d.ny * d.NK;
        hipLaunchKernelGGL(k_vort_spec<T>, dim3((n + BLK - 1) / BLK, nl * Ma), dim3(BLK), 0,
                           stream, (const cplx<T>*)sU, (const cplx<T>*)sV, (const cplx<T>*)sW,
                           (const cplx<T>*)spec[6], (const cplx<T>*)(spec[6] + d.splane), spec[3],
                           spec[4], spec[5], d.ny, d.NK, d.nz, d.splane,
                           (T)(2.0 * PI_D / d.prm.lx), (T)(2.0 * PI_D / d.prm.ly),
                           (T)(1.0 / d.dz), msp);
        ps[0] = sU;
        ps[1] = sV;
        ps[2] = sW;
    } else {
        const T* src[3] = {fld[LG_U], fld[LG_V], fld[LG_W]};
        const T scs[3] = {c, c, c};
        cplx<T>* S[6] = {spec[0], spec[1], spec[2], spec[3], spec[4], spec[5]};
        fwd2d(3, src, scs, d.plane, 1, d.field, S, d.splane, 1, msp, nl);
        VortArgs<T> va{fld[LG_DUDY], fld[LG_DUDZ], fld[LG_DVDX], fld[LG_DVDZ], fld[LG_DWDX],
                       fld[LG_DWDY], tmp[0], tmp[1], tmp[2], c};
        hipLaunchKernelGGL(k_vort<T>, grid_ij(d.nx, d.ny, nl, Ma), dim3(BLK), 0, stream, va,
                           GEO(1, nl));
        const T* vs[3] = {tmp[0], tmp[1], tmp[2]};
        fwd2d(3, vs, nullptr, d.plane, 1, d.field, S + 3, d.splane, 1, msp, nl);
    }
    pad_inv_big(6, ps, 1, nl);
    cplx<T>* cd[3] = {spec[0], spec[1], spec[2]};
    const double pc = 1.0 / (d.nx2 * d.ny2);
    if (prod_fused) {
        const ProdSrc<T> p{big[0], big[1], big[2], big[3], big[4], big[5], pc, d.nz};
        if (!launch_rows_r2c_prod<T>(timer, pbx.plan, p, d.bplane, d.ld_big, 1, mbp, bigspec,
                                     d.bsplane, mbsp, d.ny2, d.NK, nl, Ma, stream))
            throw std::runtime_error("convec: no fused products pass for nx2");
        cols_unpad(3, cd, 1, nl);
    } else {
        ProdArgs<T> pa{big[0], big[1], big[2], big[3], big[4], big[5], big[6], big[7], big[8], pc};
        const int pch = (d.nz + PROD_L - 1) / PROD_L;
        hipLaunchKernelGGL(k_products<T>, grid_ij((d.nx2 + 1) / 2, d.ny2, pch, Ma), dim3(BLK), 0,
                           stream, pa, d.nx2, d.ny2, d.ld_big, d.nz, mbp);
        const T* cs[3] = {big[6], big[7], big[8]};
        fwd_big_unpad(3, cs, cd, 1, nl);
    }
    if (spectral_out) return;          // the advection spectra stay in spec[0..2]
    const cplx<T>* is[3] = {spec[0], spec[1], spec[2]};
    const int ops[3] = {OP_NONE, OP_NONE, OP_NONE};
    T* dst[3] = {fld[LG_RHSX], fld[LG_RHSY], fld[LG_RHSZ]};
    cplx<T>* tsp[3] = {spec[3], spec[4], spec[5]};
    inv2d(3, is, ops, d.splane, 1, msp, dst, d.plane, 1, d.field, nl, tsp);
    fill_levels({{LG_RHSX, 0, BOGUS}, {LG_RHSY, 0, BOGUS}, {LG_RHSZ, 0, BOGUS},
                 {LG_RHSX, d.nz, BOGUS}, {LG_RHSY, d.nz, BOGUS}});
}

// ---------------------------------------------------------------------------
// press_stag_array.f90 + tridag_array.f90
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::press_stag_array() {
    const double cst = 1.0 / (d.nx * d.ny);
    const T c2 = (T)(cst / d.tadv1 / d.prm.dt);
    {
        // u*, v* are used on levels 1..nz-1 and w* on 1..nz: one pass over
        // 1..nz (the u*, v* level nz spectra are not read)
        // with use_cfl_dt the scale is each member's (CK_PSCALE of its clock)
        const T* src[3] = {fld[LG_U], fld[LG_V], fld[LG_W]};
        const T scs[3] = {c2, c2, c2};
        cplx<T>* S[3] = {spec[0], spec[1], spec[2]};
        fwd2d(3, src, scs, d.plane, 1, d.field, S, d.splane, 1, msp, d.nz,
              cfl_mode ? d_clock + CK_PSCALE : nullptr);
    }
    {
        // divtz at k = 1 and k = nz into planes 0 and 1 of spec[3]
        const T* src[2] = {lev(LG_DIVTZ, 1), lev(LG_DIVTZ, d.nz)};
        const T scs[2] = {(T)cst, (T)cst};
        cplx<T>* S[2] = {spec[3], spec[3] + d.splane};
        fwd2d(2, src, scs, d.plane, 0, d.field, S, d.splane, 0, msp, 1);
    }
    TriGeo tg{d.ny, d.NK, d.nz, d.splane, msp, Ma, 2.0 * PI_D / d.prm.lx,
              2.0 * PI_D / d.prm.ly, d.dz, 1.0 / (d.dz * d.dz), 1.0 / d.dz};
    const int nsys = Ma * d.ny * d.NK;
    const double tbytes =
        (double)Ma * d.splane * sizeof(cplx<T>) * (3.0 * d.nz + 1 + 2 + d.nz + 1);
    const PassChoice ch = choice_for<T>(PK_TRIDAG, d.ny, d.NK);
    const int tb = ch.S;
    timer.run(stream, KK_TRIDAG, tbytes, [&] {
        LG_LAUNCH(k_tridag<T>, dim3((nsys + tb - 1) / tb), dim3(tb), 0, stream,
                           (const cplx<T>*)spec[0], (const cplx<T>*)spec[1],
                           (const cplx<T>*)spec[2], (const cplx<T>*)spec[3],
                           (const cplx<T>*)(spec[3] + d.splane), spec[4], gam, usc, tg, dev_err);
    });
    {
        const cplx<T>* is[1] = {spec[4]};
        const int ops[1] = {OP_NONE};
        T* dst[1] = {fld[LG_P]};
        cplx<T>* tsp[1] = {spec[5]};
        inv2d(1, is, ops, d.splane, 0, msp, dst, d.plane, 0, d.field, d.nz + 1, tsp);
    }
    {
        const cplx<T>* is[2] = {spec[4], spec[4]};
        const int ops[2] = {OP_IKX, OP_IKY};
        T* dst[2] = {fld[LG_DPDX], fld[LG_DPDY]};
        cplx<T>* tsp[2] = {spec[5], spec[6]};
        inv2d(2, is, ops, d.splane, 1, msp, dst, d.plane, 1, d.field, d.nz - 1, tsp);
    }
    hipLaunchKernelGGL(k_dpdz<T>, grid_ij(d.nx, d.ny, d.nz, Ma), dim3(BLK), 0, stream,
                       (const T*)fld[LG_P], fld[LG_DPDZ], GEO(1, d.nz), (T)d.dz);
    fill_levels({{LG_DPDX, d.nz, BOGUS}, {LG_DPDY, d.nz, BOGUS}});
}

// ---------------------------------------------------------------------------
// forcing.f90::project
// ---------------------------------------------------------------------------
template <typename T>
void Case<T>::project() {
    hipLaunchKernelGGL(k_project<T>, grid_ij(d.nx, d.ny, d.nz - 1, Ma), dim3(BLK), 0, stream,
                       fld[LG_U], fld[LG_V], fld[LG_W], (const T*)fld[LG_DPDX],
                       (const T*)fld[LG_DPDY], (const T*)fld[LG_DPDZ], GEO(1, d.nz - 1),
                       (T)d.prm.dt, (T)d.tadv1, clk());
    const dim3 gp((unsigned)((d.plane + BLK - 1) / BLK), Ma);
    hipLaunchKernelGGL(k_copy_plane<T>, gp, dim3(BLK), 0, stream, (const T*)lev(LG_U, d.nz - 1),
                       lev(LG_U, d.nz), d.field, d.plane);
    hipLaunchKernelGGL(k_copy_plane<T>, gp, dim3(BLK), 0, stream, (const T*)lev(LG_V, d.nz - 1),
                       lev(LG_V, d.nz), d.field, d.plane);
    fill_levels({{LG_W, d.nz, 0.0}, {LG_W, 1, 0.0}});
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
                 tb ? fld[LG_FZA] : nullptr, (T)d.prm.dt, (T)d.tadv1, (T)d.tadv2, d_mpf,
                 (const int*)d_flags, d.prm.use_mean_p_force, clk()};
    hipLaunchKernelGGL(k_rhs_ab2<T>, grid_ij(d.ld, d.ny, d.nz, Ma), dim3(BLK), 0, stream, a,
                       GEO(1, d.nz));
    fill_levels({{LG_U, 0, BOGUS}, {LG_V, 0, BOGUS}, {LG_W, 0, BOGUS}, {LG_U, d.nz, BOGUS},
                 {LG_V, d.nz, BOGUS}});
}

template <typename T>
void Case<T>::rhs_gradp() {
    hipLaunchKernelGGL(k_rhs_gradp<T>, grid_ij(d.ld, d.ny, d.nz, Ma), dim3(BLK), 0, stream,
                       fld[LG_RHSX], fld[LG_RHSY], fld[LG_RHSZ], (const T*)fld[LG_DPDX],
                       (const T*)fld[LG_DPDY], (const T*)fld[LG_DPDZ], GEO(1, d.nz));
}

// The loop body after the counters advance and RHS_f = RHS (main.f90:155-326).
template <typename T>
void Case<T>::step_body() {
    clock_begin();
    filt_da_uvw();
    ddz_uvw();
    wallstress();
    sgs_stag();
    divstress_uvw();
    convec(true);
    if (turb) turbines_forcing();
    rhs_ab2();
    press_stag_array();
    rhs_gradp();
    project();
}


// Per-member flags of the step the counters describe: Euler update and
// Cs_opt2 = 0.03 at a fresh member's first step; dynamic-model mode (0 not
// yet, 1 update, 2 first call with F_* initialisation) when jt_total is a
// multiple of cs_count.  Writes them to the device if they changed: with at
// most FLAGS_ARG_MAX values (6 per member, M <= 128) by a k_set_flags launch
// carrying them as kernel arguments, in stream order with no host wait; for
// more members by a copy that synchronises the stream.  Steps that change
// flags run eagerly, never inside a capture.  Returns true when the step has
// a first-step or initialisation event.  force_lag: a direct lagrange_Sdep
// call, which updates every member.
template <typename T>
bool Case<T>::refresh_flags(bool force_lag) {
    const int cs = d.prm.cs_count;
    const bool dyn_step = jt_total % cs == 0;
    bool special = false, any_lag = false;
    std::fill(hflags.begin(), hflags.end(), 0);
    for (int m = 0; m < Ma; ++m) {
        int eu = 0, ci = 0, lm = 0;
        const int j = mjt[m];
        if (d.prm.initu) {
            lm = (dyn_step || force_lag) ? 1 : 0;
        } else {
            eu = ci = (j == 1) ? 1 : 0;
            if ((!ci && j >= d.prm.dyn_init && dyn_step) || force_lag) {
                lm = 1;
                if (!lm_init[m] && (j == cs || j == d.prm.dyn_init)) lm = 2;
            }
        }
        hflags[m] = eu;
        hflags[M + m] = ci;
        hflags[2 * M + m] = lm;
        hflags[5 * M + m] = (spec_on && mcold[m]) ? 1 : 0;
        if (cfl_mode) {
            hflags[3 * M + m] = clk_fresh[m] ? 1 : 0;
            hflags[4 * M + m] = (d.prm.initu || j >= d.prm.dyn_init - cs + 1) ? 1 : 0;
        }
        special |= eu || ci || lm == 2;
        any_lag |= lm > 0;
    }
    lag_now = any_lag;
    if (hflags != dflags_now) {
        if ((int)hflags.size() <= FLAGS_ARG_MAX) {
            // in stream order, from the kernel's arguments: no host wait
            FlagArgs fa;
            fa.n = (int)hflags.size();
            std::copy(hflags.begin(), hflags.end(), fa.v);
            hipLaunchKernelGGL(k_set_flags, dim3(1), dim3(256), 0, stream, d_flags, fa);
            LG_HIP(hipGetLastError());
        } else {
            LG_HIP(hipStreamSynchronize(stream));
            LG_HIP(hipMemcpy(d_flags, hflags.data(), hflags.size() * sizeof(int),
                             hipMemcpyHostToDevice));
        }
        dflags_now = hflags;
    }
    return special;
}

// Host side of the clocks for the step the flags describe: total_time of a
// fixed dt (main.f90:149); a CFL clock is fresh for one step only.
template <typename T>
void Case<T>::advance_clocks() {
    for (int m = 0; m < Ma; ++m) {
        if (!cfl_mode) mtime[m] = mtime[m] + d.prm.dt;
        clk_fresh[m] = 0;
    }
}

// ---------------------------------------------------------------------------
// Spectral-resident step (spectral.hpp)
// ---------------------------------------------------------------------------
template <typename T>
static SpecGeo<T> spec_geo(const Dims& d, long msp, int k0, int nlev) {
    return SpecGeo<T>{d.ny, d.NK, d.nz, k0, nlev, d.splane, msp, (T)(2.0 * PI_D / d.prm.lx),
                      (T)(2.0 * PI_D / d.prm.ly)};
}

// The six stress spectra (levels 1..nz) for divstress_uv + divstress_w,
// which k_spec_ab2 forms from them; no inverse transform.
template <typename T>
void Case<T>::divstress_spec() {
    // the stress spectra go to spec[3..8] (convec has finished with them;
    // its advection spectra stay in spec[0..2]); k_spec_ab2 forms divt
    const T sc = (T)(1.0 / (d.nx * d.ny));
    const T* src[6] = {fld[LG_TXX], fld[LG_TXY], fld[LG_TYY],
                       fld[LG_TXZ], fld[LG_TYZ], fld[LG_TZZ]};
    const T scs[6] = {sc, sc, sc, sc, sc, sc};
    cplx<T>* S[6] = {spec[3], spec[4], spec[5], spec[6], spec[7], spec[8]};
    fwd2d(6, src, scs, d.plane, 1, d.field, S, d.splane, 1, msp, d.nz);
}

// press_stag_array on u*, v*, w* spectra (in sU, sV, sW; the solve scales
// them by 1/(tadv1 dt)) and the divtz spectra of levels 1 and nz; p_hat
// into sP.
template <typename T>
void Case<T>::press_spec() {
    TriGeo tg{d.ny, d.NK, d.nz, d.splane, msp, Ma, 2.0 * PI_D / d.prm.lx,
              2.0 * PI_D / d.prm.ly, d.dz, 1.0 / (d.dz * d.dz), 1.0 / d.dz};
    tg.hsc = (1.0 / d.tadv1) / d.prm.dt;
    const int nsys = Ma * d.ny * d.NK;
    const double tbytes =
        (double)Ma * d.splane * sizeof(cplx<T>) * (3.0 * d.nz + 1 + 2 + d.nz + 1);
    const PassChoice ch = choice_for<T>(PK_TRIDAG, d.ny, d.NK);
    const int tb = ch.S;
    timer.run(stream, KK_TRIDAG, tbytes, [&] {
        LG_LAUNCH(k_tridag<T>, dim3((nsys + tb - 1) / tb), dim3(tb), 0, stream,
                  (const cplx<T>*)sU, (const cplx<T>*)sV, (const cplx<T>*)sW,
                  (const cplx<T>*)(sDz + d.splane), (const cplx<T>*)(sDz + d.nz * d.splane),
                  sP, gam, usc, tg, dev_err);
    });
}

// The loop body with the state in spectra.  cold: u, v, w and RHS_f are the
// physical fields (their spectra are made first); warm: sU, sV, sW and sRf
// hold them from the previous step.
template <typename T>
void Case<T>::step_body_spec(bool cold) {
    const T sc = (T)(1.0 / (d.nx * d.ny));
    if (cold) {
        // spectra of the cold members' physical u, v, w (levels 0..nz) and
        // RHS_f (x, y on 1..nz-1, z on 1..nz): straight into the state when
        // every member is cold, else into scratch and then per member
        const bool all = count_cold() == Ma;
        const T* us[3] = {fld[LG_U], fld[LG_V], fld[LG_W]};
        const T scs[3] = {sc, sc, sc};
        cplx<T>* ud[3] = {sU, sV, sW};
        cplx<T>* rd[3] = {sRf[0], sRf[1], sRf[2]};
        if (!all) {
            for (int q = 0; q < 3; ++q) {
                ud[q] = spec[q];
                rd[q] = spec[3 + q];
            }
        }
        fwd2d(3, us, scs, d.plane, 0, d.field, ud, d.splane, 0, msp, d.nz + 1);
        const T* rs[2] = {fld[LG_RHSX_F], fld[LG_RHSY_F]};
        fwd2d(2, rs, scs, d.plane, 1, d.field, rd, d.splane, 1, msp, d.nz - 1);
        const T* rz[1] = {fld[LG_RHSZ_F]};
        fwd2d(1, rz, &sc, d.plane, 1, d.field, rd + 2, d.splane, 1, msp, d.nz);
        if (!all) {
            SelArgs<T> sa{};
            cplx<T>* dsts[6] = {sU, sV, sW, sRf[0], sRf[1], sRf[2]};
            for (int q = 0; q < 6; ++q) {
                sa.src[q] = spec[q];
                sa.dst[q] = dsts[q];
            }
            const long n = (long)(d.nz + 1) * d.splane;
            hipLaunchKernelGGL(k_select_members<T>, dim3((unsigned)((n + BLK - 1) / BLK), Ma, 6),
                               dim3(BLK), 0, stream, sa, (const int*)(d_flags + 5 * M), msp, n);
        }
    }
    filt_da_uvw(false);
    ddz_uvw();
    wallstress(sU, sV);
    sgs_stag();
    convec(true, true);
    divstress_spec();
    const bool tb = turb && turb->total_nodes > 0 && kf1 >= kf0;
    if (tb) {
        turbines_forcing();
        const T* fs[3] = {fld[LG_FXA], fld[LG_FYA], fld[LG_FZA]};
        const T scs[3] = {sc, sc, sc};
        fwd2d(3, fs, scs, d.plane, kf0, d.field, sF, d.splane, kf0, msp, kf1 - kf0 + 1);
    }
    const int n = d.ny * d.NK;
    const dim3 gs((n + BLK - 1) / BLK, d.nz * Ma);
    const SpecGeo<T> g = spec_geo<T>(d, msp, 1, d.nz);
    SpecAb2<T> a{spec[0], spec[1], spec[2],
                 StressSpec<T>{spec[3], spec[4], spec[5], spec[6], spec[7], spec[8]}, sDz,
                 (T)(1.0 / d.dz),
                 tb ? sF[0] : nullptr, tb ? sF[1] : nullptr, tb ? sF[2] : nullptr,
                 sR[0], sR[1], sR[2], sRf[0], sRf[1], sRf[2], sU, sV, sW, d_mpf,
                 (const int*)d_flags, d.prm.use_mean_p_force, kf0, kf1,
                 (T)d.prm.dt, (T)d.tadv1, (T)d.tadv2};
    hipLaunchKernelGGL(k_spec_ab2<T>, gs, dim3(BLK), 0, stream, a, g);
    press_spec();
    hipLaunchKernelGGL(k_spec_project<T>, gs, dim3(BLK), 0, stream, (const cplx<T>*)sP, sR[0],
                       sR[1], sR[2], sU, sV, sW, g, (T)d.prm.dt, (T)d.tadv1, (T)d.dz);
    hipLaunchKernelGGL(k_spec_bc<T>, dim3((n + BLK - 1) / BLK, Ma), dim3(BLK), 0, stream, sU,
                       sV, sW, g);
}

// The storage exchange of a step (RHS_f = RHS): physical and spectral RHS,
// and in spectral mode the disk forcing of the previous step; the staleness
// of RHS moves to RHS_f.
template <typename T>
void Case<T>::swap_rhs() {
    std::swap(fld[LG_RHSX], fld[LG_RHSX_F]);
    std::swap(fld[LG_RHSY], fld[LG_RHSY_F]);
    std::swap(fld[LG_RHSZ], fld[LG_RHSZ_F]);
    swapped = !swapped;
    if (!spec_on) return;
    for (int q = 0; q < 3; ++q) std::swap(sR[q], sRf[q]);
    std::swap(fld[LG_FXA], fprev[0]);
    std::swap(fld[LG_FYA], fprev[1]);
    std::swap(fld[LG_FZA], fprev[2]);
    std::swap(fld[LG_TXZ], tprev[0]);
    std::swap(fld[LG_TYZ], tprev[1]);
    std::swap(fld[LG_TZZ], tprev[2]);
    const bool rhs_stale = (stale & SG_RHS) != 0;
    stale = (stale & ~(unsigned)(SG_RHS | SG_RHSF)) | (rhs_stale ? (unsigned)SG_RHSF : 0u);
    for (int m = 0; m < M; ++m) rhsf_phys[m] = mcold[m];
    forced_prev = forced_last;
    forced_last = turb && turb->total_nodes > 0 && kf1 >= kf0;
    // the buffer now written holds the forcing from before a disk-set
    // change (zero off the old nodes, not the new ones): cleared once the
    // step it fed has been made physical and left behind
    if (fclear_pending > 0 && --fclear_pending == 0)
        for (int id : {LG_FXA, LG_FYA, LG_FZA})
            LG_HIP(hipMemsetAsync(fld[id], 0, M * d.field * sizeof(T), stream));
}

template <typename T>
void Case<T>::run_body(bool cold) {
    if (spec_on) step_body_spec(cold);
    else step_body();
}

template <typename T>
void Case<T>::after_body() {
    if (spec_on) {
        // a step that started from physical fields leaves RHS_f as it was,
        // except for Euler members (RHS_f = RHS)
        const bool cold = count_cold() > 0;
        bool any_euler = false;
        for (int m = 0; m < Ma; ++m) any_euler |= hflags[m] != 0;
        stale |= SG_UVW | SG_RHS | SG_P | SG_DIVT;
        if (cold && any_euler) stale |= SG_RHSF;
        // S_ij and Nu_t are stored only by a step with a dynamic-model update
        // (set here too: a replayed graph runs no host code of sgs_stag)
        if (lag_now) stale &= ~(unsigned)SG_SIJ;
        else stale |= SG_SIJ;
        std::fill(mcold.begin(), mcold.end(), 0);
    }
    if (st.on) stats_after_step();
}

// f - (its band-limited part): out = f - F^-1(trunc F(f)) on levels ka..kb.
template <typename T>
void Case<T>::high_part(T* const* f, T* const* out, int nf, int ka, int kb) {
    const int nl = kb - ka + 1;
    if (nl <= 0) return;
    const T sc = (T)(1.0 / (d.nx * d.ny));
    const T scs[3] = {sc, sc, sc};
    const T* fs[3] = {f[0], f[1], f[2]};
    cplx<T>* S[3] = {spec[0], spec[1], spec[2]};
    fwd2d(nf, fs, scs, d.plane, ka, d.field, S, d.splane, ka, msp, nl);
    const cplx<T>* is[3] = {spec[0], spec[1], spec[2]};
    const int ops[3] = {OP_NONE, OP_NONE, OP_NONE};
    cplx<T>* tsp[3] = {spec[6], spec[7], spec[8]};
    inv2d(nf, is, ops, d.splane, ka, msp, out, d.plane, ka, d.field, nl, tsp);
    for (int q = 0; q < nf; ++q)
        hipLaunchKernelGGL(k_sub<T>, grid_ij(d.ld, d.ny, nl, Ma), dim3(BLK), 0, stream, out[q],
                           (const T*)f[q], (const T*)out[q], GEO(ka, nl));
}

// The part of a step's RHS that its spectra do not carry: the
// high-wavenumber (Nyquist) content of the disk forcing and of the vertical
// stress derivatives, which the reference's physical RHS holds and its next
// filt_da removes.  g = hp(f) - d/dz hp(t) with hp = x - F^-1(trunc F(x)):
// x: hp(fx) - (hp(txz)(k+1) - hp(txz)(k))/dz, y alike with tyz, z: hp(fz) -
// (hp(tzz)(k) - hp(tzz)(k-1))/dz on 2..nz-1; levels 1..nz-1.  f may be null
// (no disks).  scratch: three fields.
template <typename T>
void Case<T>::rhs_high(T* const* f, T* const* t, T* const* g, T* const* scratch) {
    const int nz = d.nz, n1 = nz - 1;
    high_part(t, scratch, 3, 1, nz);
    if (f) high_part(f, g, 3, 1, n1);
    hipLaunchKernelGGL(k_rhs_high<T>, grid_ij(d.ld, d.ny, n1, Ma), dim3(BLK), 0, stream, g[0],
                       g[1], g[2], (const T*)scratch[0], (const T*)scratch[1],
                       (const T*)scratch[2], f != nullptr, GEO(1, n1), (T)(1.0 / d.dz));
}

// Make the stale physical fields of `groups` from the spectra, with the
// high-wavenumber remainders (rhs_high) added back where the reference's
// fields hold them.
template <typename T>
void Case<T>::materialize(unsigned groups) {
    groups &= stale;
    if (!groups) return;
    const int nz = d.nz, n1 = nz - 1;
    // gc: this step's remainder; gq: the previous step's (or the high part of
    // a physical RHS_f); hs: scratch of rhs_high
    T* gc[3] = {tmp[0], tmp[1], tmp[2]};
    T* gq[3] = {tmp[3], tmp[4], tmp[5]};
    T* hs[3] = {tmp[6], tmp[7], tmp[8]};
    bool have_gq = false;
    if (groups & SG_DIVT) {
        // the step keeps no divt spectra (k_spec_ab2 forms divt from the
        // stress spectra): the reference's divstress_uv/w on the step's
        // physical stresses (uses spec[0..7] and tmp[0], tmp[2], tmp[4],
        // before the remainders below take tmp)
        divstress_uvw();
    }
    const unsigned need_g = SG_UVW | SG_RHS | SG_RHSF;
    if (groups & need_g) {
        T* f[3] = {fld[LG_FXA], fld[LG_FYA], fld[LG_FZA]};
        T* t[3] = {fld[LG_TXZ], fld[LG_TYZ], fld[LG_TZZ]};
        rhs_high(forced_last ? f : nullptr, t, gc, hs);
    }
    if (groups & (SG_UVW | SG_RHSF)) {
        // per member: the step before's stresses and forcing, or, for a
        // member whose last step started from its physical fields, the high
        // part of its physical RHS_f (selected by the step's cold flags)
        int nphys = 0;
        for (int m = 0; m < Ma; ++m) nphys += rhsf_phys[m] ? 1 : 0;
        if (nphys < Ma) {
            T* t[3] = {tprev[0], tprev[1], tprev[2]};
            rhs_high(forced_prev ? fprev : nullptr, t, gq, hs);
        }
        if (nphys > 0) {
            T* f[3] = {fld[LG_RHSX_F], fld[LG_RHSY_F], fld[LG_RHSZ_F]};
            high_part(f, nphys == Ma ? gq : hs, 3, 1, n1);
            if (nphys < Ma) {
                // whole fields as complex pairs (d.field is even)
                SelArgs<T> sa{};
                for (int q = 0; q < 3; ++q) {
                    sa.src[q] = reinterpret_cast<const cplx<T>*>(hs[q]);
                    sa.dst[q] = reinterpret_cast<cplx<T>*>(gq[q]);
                }
                const long n = d.field / 2;
                hipLaunchKernelGGL(k_select_members<T>,
                                   dim3((unsigned)((n + BLK - 1) / BLK), Ma, 3), dim3(BLK), 0,
                                   stream, sa, (const int*)(d_flags + 5 * M), n, n);
            }
        }
        have_gq = true;
    }
    const int on[3] = {OP_NONE, OP_NONE, OP_NONE};
    cplx<T>* tsp[3] = {spec[9], spec[10], spec[11]};
    const T dt = (T)d.prm.dt, t1 = (T)d.tadv1, t2 = (T)d.tadv2;
    if (groups & SG_UVW) {
        const cplx<T>* is[3] = {sU, sV, sW};
        T* dst[3] = {fld[LG_U], fld[LG_V], fld[LG_W]};
        inv2d(3, is, on, d.splane, 1, msp, dst, d.plane, 1, d.field, nz, tsp);
        // u = F^-1(u_hat) + dt (t1 gc + t2 (Euler ? gc : gq))
        for (int q = 0; q < 3; ++q)
            hipLaunchKernelGGL(k_add_high<T>, grid_ij(d.ld, d.ny, n1, Ma), dim3(BLK), 0, stream,
                               fld[LG_U + q], (const T*)gc[q], have_gq ? (const T*)gq[q] : nullptr,
                               (const int*)d_flags, GEO(1, n1), dt, t1, t2);
        const dim3 gp((unsigned)((d.plane + BLK - 1) / BLK), Ma);
        hipLaunchKernelGGL(k_copy_plane<T>, gp, dim3(BLK), 0, stream,
                           (const T*)lev(LG_U, nz - 1), lev(LG_U, nz), d.field, d.plane);
        hipLaunchKernelGGL(k_copy_plane<T>, gp, dim3(BLK), 0, stream,
                           (const T*)lev(LG_V, nz - 1), lev(LG_V, nz), d.field, d.plane);
        fill_levels({{LG_W, nz, 0.0}, {LG_W, 1, 0.0}, {LG_U, 0, BOGUS}, {LG_V, 0, BOGUS},
                     {LG_W, 0, BOGUS}});
    }
    if (groups & SG_RHS) {
        const cplx<T>* is[3] = {sR[0], sR[1], sR[2]};
        T* dst[3] = {fld[LG_RHSX], fld[LG_RHSY], fld[LG_RHSZ]};
        inv2d(3, is, on, d.splane, 1, msp, dst, d.plane, 1, d.field, nz, tsp);
        for (int q = 0; q < 3; ++q)
            hipLaunchKernelGGL(k_add<T>, grid_ij(d.ld, d.ny, n1, Ma), dim3(BLK), 0, stream,
                               fld[LG_RHSX + q], (const T*)gc[q], GEO(1, n1));
        fill_levels({{LG_RHSX, 0, BOGUS}, {LG_RHSY, 0, BOGUS}, {LG_RHSZ, 0, BOGUS},
                     {LG_RHSX, nz, BOGUS}, {LG_RHSY, nz, BOGUS}});
    }
    if (groups & SG_RHSF) {
        // RHS_f = the previous RHS (or the physical input of a step that
        // started from it), or this step's (before - grad p) for an Euler
        // member
        const cplx<T>* is[3] = {sRf[0], sRf[1], sRf[2]};
        T* dst[3] = {fld[LG_RHSX_F], fld[LG_RHSY_F], fld[LG_RHSZ_F]};
        inv2d(3, is, on, d.splane, 1, msp, dst, d.plane, 1, d.field, nz, tsp);
        for (int q = 0; q < 3; ++q)
            hipLaunchKernelGGL(k_add_high<T>, grid_ij(d.ld, d.ny, n1, Ma), dim3(BLK), 0, stream,
                               fld[LG_RHSX_F + q], (const T*)gc[q], (const T*)gq[q],
                               (const int*)d_flags, GEO(1, n1), (T)1, (T)0, (T)1);
        fill_levels({{LG_RHSX_F, 0, BOGUS}, {LG_RHSY_F, 0, BOGUS}, {LG_RHSZ_F, 0, BOGUS},
                     {LG_RHSX_F, nz, BOGUS}, {LG_RHSY_F, nz, BOGUS}});
    }
    if (groups & SG_P) {
        const cplx<T>* is[1] = {sP};
        T* dst[1] = {fld[LG_P]};
        inv2d(1, is, on, d.splane, 0, msp, dst, d.plane, 0, d.field, nz + 1, tsp);
        const cplx<T>* is2[2] = {sP, sP};
        const int ops[2] = {OP_IKX, OP_IKY};
        T* dst2[2] = {fld[LG_DPDX], fld[LG_DPDY]};
        inv2d(2, is2, ops, d.splane, 1, msp, dst2, d.plane, 1, d.field, n1, tsp);
        hipLaunchKernelGGL(k_dpdz<T>, grid_ij(d.nx, d.ny, nz, Ma), dim3(BLK), 0, stream,
                           (const T*)fld[LG_P], fld[LG_DPDZ], GEO(1, nz), (T)d.dz);
        fill_levels({{LG_DPDX, nz, BOGUS}, {LG_DPDY, nz, BOGUS}});
    }
    if (groups & SG_SIJ) {
        calc_sij();
        TauArgs<T> a{fld[LG_S11], fld[LG_S12], fld[LG_S13], fld[LG_S22], fld[LG_S23],
                     fld[LG_S33], fld[LG_CS_OPT2], fld[LG_NU_T], nullptr, nullptr, nullptr,
                     nullptr, nullptr, nullptr, (T)(d.delta * d.delta), (T)d.prm.nu};
        hipLaunchKernelGGL(k_nut<T>, grid_ij(d.nx, d.ny, d.nz, Ma), dim3(BLK), 0, stream, a,
                           GEO(1, d.nz));
    }
    stale &= ~groups;
    LG_HIP(hipGetLastError());
}

// Before the host reads (modify = false) or changes (true) physical
// fields: make the stale ones; a change also makes the next step start
// from the physical fields.
template <typename T>
void Case<T>::touch(bool modify) {
    if (!spec_on) return;
    materialize(SG_ALL);
    if (modify) std::fill(mcold.begin(), mcold.end(), 1);
}

template <typename T>
int Case<T>::count_cold() const {
    int n = 0;
    for (int m = 0; m < Ma; ++m) n += mcold[m] ? 1 : 0;
    return n;
}

// Member dst takes member src's spectral state (with its physical fields
// copied by the caller, the two members step alike).
template <typename T>
void Case<T>::copy_spec_member(int src, int dst) {
    if (!spec_on) return;
    cplx<T>* bufs[10] = {sU, sV, sW, sR[0], sR[1], sR[2], sRf[0], sRf[1], sRf[2], sP};
    for (cplx<T>* b : bufs)
        LG_HIP(hipMemcpyAsync(b + (long)dst * msp, b + (long)src * msp, msp * sizeof(cplx<T>),
                              hipMemcpyDeviceToDevice, stream));
    for (int q = 0; q < 3; ++q) {
        LG_HIP(hipMemcpyAsync(fprev[q] + (long)dst * d.field, fprev[q] + (long)src * d.field,
                              d.field * sizeof(T), hipMemcpyDeviceToDevice, stream));
        LG_HIP(hipMemcpyAsync(tprev[q] + (long)dst * d.field, tprev[q] + (long)src * d.field,
                              d.field * sizeof(T), hipMemcpyDeviceToDevice, stream));
    }
    mcold[dst] = mcold[src];
}

template <typename T>
void Case<T>::step() {
    ++jt;
    ++jt_total;
    for (int m = 0; m < Ma; ++m) ++mjt[m];
    refresh_flags();
    advance_clocks();
    const bool cold = spec_on && count_cold() > 0;
    swap_rhs();
    run_body(cold);
    after_body();
}

template <typename T>
void Case<T>::drop_graphs() {
    for (int q = 0; q < 4; ++q) {
        if (graphs[q]) (void)hipGraphExecDestroy(graphs[q]);
        graphs[q] = nullptr;
        graph_ma[q] = 0;
    }
}

// One step replayed from a captured graph.  A step's launches depend on the
// RHS storage parity and on whether the dynamic model runs, so up to four
// graphs are captured, each on first use (and again after the member count
// or the disk set changes).  A member's first step, its first dynamic call
// and a step that starts from physical fields run eagerly.
template <typename T>
void Case<T>::step_graph() {
    ++jt;
    ++jt_total;
    for (int m = 0; m < Ma; ++m) ++mjt[m];
    const bool special = refresh_flags();
    advance_clocks();
    const bool cold = spec_on && count_cold() > 0;
    swap_rhs();
    if (special || cold) {
        run_body(cold);
        after_body();
        return;
    }
    const int key = (swapped ? 1 : 0) + (lag_now ? 2 : 0);
    if (!graphs[key] || graph_ma[key] != Ma) {
        if (graphs[key]) (void)hipGraphExecDestroy(graphs[key]);
        graphs[key] = nullptr;
        const bool ton = timer.on;
        timer.on = false;
        LG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            run_body(false);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            timer.on = ton;
            throw;
        }
        timer.on = ton;
        LG_HIP(hipStreamEndCapture(stream, &g));
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        LG_HIP(e);
        graphs[key] = ex;
        graph_ma[key] = Ma;
    }
    LG_HIP(hipGraphLaunch(graphs[key], stream));
    after_body();
}

// ---------------------------------------------------------------------------
// Diagnostics (member m)
// ---------------------------------------------------------------------------
template <typename T>
double Case<T>::rmsdiv(int m) {
    const int nb = 256;
    hipLaunchKernelGGL(k_rmsdiv_part<T>, dim3(nb), dim3(BLK), 0, stream,
                       (const T*)mem(LG_DUDX, m), (const T*)mem(LG_DVDY, m),
                       (const T*)mem(LG_DWDZ, m), GEO(1, d.nz), red);
    std::vector<double> h(nb);
    LG_HIP(hipMemcpyAsync(h.data(), red, nb * sizeof(double), hipMemcpyDeviceToHost, stream));
    LG_HIP(hipStreamSynchronize(stream));
    double s = 0.0;
    for (double x : h) s += x;
    return s / ((double)d.nx * d.ny * (d.nz - 1));
}

template <typename T>
double Case<T>::energy(int m) {
    touch(false);
    const int nb = 256;
    hipLaunchKernelGGL(k_energy_part<T>, dim3(nb), dim3(BLK), 0, stream, (const T*)mem(LG_U, m),
                       (const T*)mem(LG_V, m), (const T*)mem(LG_W, m), GEO(1, d.nz), red);
    std::vector<double> h(nb);
    LG_HIP(hipMemcpyAsync(h.data(), red, nb * sizeof(double), hipMemcpyDeviceToHost, stream));
    LG_HIP(hipStreamSynchronize(stream));
    double s = 0.0;
    for (double x : h) s += x;
    return s * 0.5 / ((double)d.nx * d.ny * (d.nz - 1));
}

template <typename T>
double Case<T>::max_cfl(int m) {
    touch(false);
    const int nb = 256;
    double mx3[3];
    const double h3[3] = {d.dx, d.dy, d.dz};
    const int ids[3] = {LG_U, LG_V, LG_W};
    for (int q = 0; q < 3; ++q) {
        hipLaunchKernelGGL(k_maxabs_part<T>, dim3(nb), dim3(BLK), 0, stream,
                           (const T*)mem(ids[q], m), GEO(1, d.nz), red);
        std::vector<double> h(nb);
        LG_HIP(hipMemcpyAsync(h.data(), red, nb * sizeof(double), hipMemcpyDeviceToHost,
                              stream));
        LG_HIP(hipStreamSynchronize(stream));
        double mx = 0.0;
        for (double x : h) mx = std::max(mx, x);
        mx3[q] = mx / h3[q];
    }
    double ck[5];
    clock_get(m, ck);
    return ck[0] * std::max(mx3[0], std::max(mx3[1], mx3[2]));
}

template <typename T>
double Case<T>::tau_wall_bot(int m) {
    std::vector<T> hx(d.plane), hy(d.plane);
    LG_HIP(hipMemcpyAsync(hx.data(), mem(LG_TXZ, m) + d.plane, d.plane * sizeof(T),
                          hipMemcpyDeviceToHost, stream));
    LG_HIP(hipMemcpyAsync(hy.data(), mem(LG_TYZ, m) + d.plane, d.plane * sizeof(T),
                          hipMemcpyDeviceToHost, stream));
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
