// Device-resident state of one LES case and the subsystem drivers.
#pragma once

#include <initializer_list>
#include <memory>
#include <vector>

#include "../../include/lesgo_amd.h"
#include "common.hpp"
#include "fft.hpp"
#include "passes.hpp"
#include "timer.hpp"

namespace lg {

// Run parameters plus the derived sizes of the one-rank layout
// (input_util.f90:196-235, 403-409; sgs_param.f90:187).
struct Dims {
    lg_params prm;
    int nx, ny, nz, lh, ld, nx2, ny2, ld_big, NK;
    double dx, dy, dz, tadv1, tadv2, delta, lagran_dt;
    long plane;       // ny * ld
    long field;       // (nz+1) * ny * ld
    long splane;      // ny * NK (complex)
    long bplane;      // ny2 * ld_big
    long bsplane;     // ny2 * NK (complex)
    explicit Dims(const lg_params& p);
};

// Device twiddle / permutation tables for one transform length.
struct PlanTables {
    FftPlan plan;
    int* perm = nullptr;
    double2* tw64 = nullptr;
    float2* tw32 = nullptr;
    PlanTables() = default;
    PlanTables(const PlanTables&) = delete;
    PlanTables& operator=(const PlanTables&) = delete;
    ~PlanTables();
    void build(int N);
};

// Actuator disks of one member (turbines.f90:275-462 output, host side).
struct MemberTurbines {
    std::vector<double> spec;        // 8 per turbine
    std::vector<double> M;           // indicator M per turbine
    std::vector<long> lin;           // node offset (k*ny + j)*ld + i within the member
    std::vector<double> ind;         // normalised indicator weight
    std::vector<int> off;            // n+1 offsets into lin / ind
    std::vector<double> coef;        // per turbine TC_N: n1,n2,n3,ct',dia,corrfac,member,-
    std::vector<double> state;       // per turbine: u_d, u_d_T, f_n (last download)
    bool adm = true;                 // ADM correction (turbines.f90:578-581)
    int n() const { return (int)spec.size() / 8; }
};
constexpr int TC_N = 8;              // coefficients per turbine

// Device node lists of every member's disks, concatenated member by member.
struct TurbineSet {
    int n = 0;                       // turbines over all members
    int total_nodes = 0;
    std::vector<int> first;          // first turbine of each member (M+1)
    long* node_lin = nullptr;        // m*field + (k*ny + j)*ld + i
    double* node_ind = nullptr;      // normalised indicator weight
    int* off = nullptr;              // n+1 offsets
    int* node_tid = nullptr;         // turbine of each node
    bool has_fz = false;
    // with tilted or yawed-up disks: fza on the uv grid before its move to
    // the w grid, one field per member, zero off the nodes (its own buffer:
    // the step's scratch fields are reused between steps)
    void* fz_uv = nullptr;
    double* coef = nullptr;          // per turbine: TC_N coefficients
    double* state = nullptr;         // per turbine: u_d, u_d_T, f_n
    long* center_lin = nullptr;      // per turbine: the centre cell, and the
    long* center_wlin = nullptr;     //   w level below its w_uv (turbines.f90:196-198)
    double* center = nullptr;        // per turbine: u, v, w_uv at the centre (last forcing)
    ~TurbineSet();
};

// Batch of M same-grid members.  Member m of every array starts at
// m * (its member stride); Ma <= M members are stepped.
template <typename T>
struct Case {
    Dims d;
    int M = 1;                       // member capacity
    int Ma = 1;                      // members stepped
    long msp = 0;                    // (nz+1) * splane
    long mbp = 0;                    // (nz+1) * bplane
    long mbsp = 0;                   // (nz+1) * bsplane
    hipStream_t stream = nullptr;
    int jt = 0, jt_total = 0;
    bool swapped = false;            // RHS / RHS_f storage exchanged (odd steps)
    std::vector<double> zo, mpf;     // per member z0 and mean_p_force (x, y)
    double* d_denom = nullptr;       // per member ln(0.5 dz / z0)
    double* d_mpf = nullptr;         // per member mean_p_force_x, _y
    std::vector<MemberTurbines> mturb;
    // fresh-start bookkeeping per member (lg_params.initu = 0)
    std::vector<int> mjt;            // steps the member has taken (its jt)
    std::vector<char> lm_init;       // F_* initialised by the first dynamic call
    // euler[M], csinit[M], lagmode[M], clock fresh[M], lagran_dt grows[M]
    // of this step
    std::vector<int> hflags;
    std::vector<int> dflags_now;     // the copy on the device
    int* d_flags = nullptr;
    bool lag_now = false;            // the dynamic model runs this step
    // time-stepping clock.  use_cfl_dt: CK_N doubles per member on the
    // device (clock.hpp), advanced by the step itself; fixed dt: only
    // total_time changes, kept on the host.
    bool cfl_mode = false;
    double* d_clock = nullptr;
    unsigned long long* d_cmax = nullptr;  // 3 per member: max |u|, |v|, |w|
    std::vector<double> mtime;       // total_time per member (fixed dt)
    std::vector<char> clk_fresh;     // the member's clock holds no previous step
    // spectral-resident stepping (spectral.hpp): with a fixed dt the step
    // keeps u, v, w, RHS, RHS_f, p and divt as spectra between steps and
    // makes the physical fields on demand (materialize).  mcold[m]: member
    // m's next step starts from its physical fields (its spectra do not
    // describe it); stale: the physical groups not yet made.
    // SG_SIJ: S_ij and Nu_t, which a step without a dynamic-model update
    // does not store (only the host reads them then; made from the
    // derivative fields, which stay valid until the next step)
    enum : unsigned {
        SG_UVW = 1, SG_RHS = 2, SG_RHSF = 4, SG_P = 8, SG_DIVT = 16, SG_SIJ = 32, SG_ALL = 63
    };
    bool spec_on = false;
    // the Lagrangian model's inverse row pass forms the tensor contractions
    // itself (k_lag_rows) where the grid allows it; LESGO_AMD_LAG_ROWS=0
    // keeps the 42 filtered fields
    bool lag_fused = true;
    // the convective products are formed in the 3/2 grid's forward row pass
    // (k_rows_r2c_prod) where that length has a four-step split;
    // LESGO_AMD_PROD_ROWS=0 keeps k_products and the three product fields
    bool prod_fused = true;
    std::vector<char> mcold;
    // members whose last step started from their physical fields (their
    // RHS_f high part is in the physical RHS_f only); the others take it
    // from the step before's stresses and forcing (tprev, fprev), so what a
    // member's fields read back as never depends on earlier reads
    std::vector<char> rhsf_phys;
    bool forced_last = false;        // the last step wrote disk forcing
    bool forced_prev = false;        // the step before it did
    // steps until the forcing buffer that held the forcing of the step
    // before a disk-set change becomes the write target and is cleared
    int fclear_pending = 0;
    unsigned stale = 0;
    cplx<T>* sR[3] = {};             // RHS spectra (after - grad p)
    cplx<T>* sRf[3] = {};            // RHS_f spectra
    cplx<T>* sP = nullptr;           // pressure spectrum, levels 0..nz
    cplx<T>* sDz = nullptr;          // divtz spectra of the step (levels 1 and nz only)
    cplx<T>* sF[3] = {};             // forcing spectra of the step (disk levels)
    T* fprev[3] = {};                // the previous step's fxa, fya, fza
    T* tprev[3] = {};                // the previous step's txz, tyz, tzz
    int kf0 = 1, kf1 = 0;            // levels where the disks force
    // running time averages (time_average.f90) and streamwise spectra
    struct StatsCfg {
        int nstart = 0, nend = -1, nskip = 1;
        bool on = false, initialized = false;
        double dt = 0.0, total_time = 0.0;
    };
    StatsCfg st;
    double* tavg = nullptr;          // 26 quantities x M x (nz+1) x ny x nx, FP64
    double* tspec = nullptr;         // 3 components x M x (nz+1) x NK, FP64
    hipGraphExec_t graphs[4] = {};   // captured steps keyed by (swapped, dynamic model)
    int graph_ma[4] = {};

    T* fld[LG_NFIELDS] = {};
    // spectral scratch, each (nz+1) small spectral planes
    static constexpr int NSPEC = 12;
    cplx<T>* spec[NSPEC] = {};
    // big-grid scratch (levels 0..nz)
    static constexpr int NBIG = 9;
    T* big[NBIG] = {};
    cplx<T>* bigspec[6] = {};
    // physical scratch fields
    static constexpr int NTMP = 9;
    T* tmp[NTMP] = {};
    // pressure solve
    cplx<T>* sU = nullptr;           // filtered u, v, w spectra of the current step
    cplx<T>* sV = nullptr;
    cplx<T>* sW = nullptr;
    double* gam = nullptr;           // (nz+1) * ny * NK: the pressure systems' gam table
    double2* usc = nullptr;          // (nz+1) * ny * NK
    int* dev_err = nullptr;
    // Lagrangian model scratch (chunked over levels)
    int lag_chunk = 0;
    T* lag_phys = nullptr;           // 15 * chunk planes
    cplx<T>* lag_spec = nullptr;     // 21 * chunk planes
    T* lag_filt = nullptr;           // 42 * chunk planes
    T* Ftmp = nullptr;               // 4 fields (interpolated F_*, or F_* copies)
    double* red = nullptr;           // reduction scratch
    T* G1 = nullptr;                 // [ny][NK]
    T* G2 = nullptr;
    // the filters' bands: columns kx < g_nzc[f] and rows |ky index| <
    // g_ykeep[f] hold every nonzero entry of G1 (f = 0) and G2 (f = 1)
    int g_nzc[2] = {0, 0}, g_ykeep[2] = {0, 0};
    PlanTables px, py, pbx, pby;     // nx, ny, nx2, ny2
    std::unique_ptr<TurbineSet> turb;
    // u_d_T time filter (turbines.f90:561-566): T_avg_dim <= 0 is none
    double t_avg_dim = -1.0, z_i = 1.0, u_star = 1.0;
    KernelTimer timer;

    Case(const lg_params& p, int members);
    ~Case();
    Case(const Case&) = delete;
    Case& operator=(const Case&) = delete;

    T* F(int id) { return fld[id]; }
    T* lev(int id, int k) { return fld[id] + (long)k * d.plane; }
    T* mem(int id, int m) { return fld[id] + (long)m * d.field; }

    // 2-D transform helpers over planes [lev0, lev0+nlev) of every member
    // band (nzc, ykeep; 0: none): the spectra are only read inside the test
    // filters' band, so only that part is computed and written
    void fwd2d(int njobs, const T* const* src, const T* scales, long src_plane, int src_lev0,
               long src_ms, cplx<T>* const* dst, long dst_plane, int dst_lev0, long dst_ms,
               int nlev, const double* mscale = nullptr, int band_nzc = 0, int band_ykeep = 0);
    void inv2d(int njobs, const cplx<T>* const* src, const int* ops, long src_plane, int src_lev0,
               long src_ms, T* const* dst, long dst_plane, int dst_lev0, long dst_ms, int nlev,
               cplx<T>* const* tmpspec, const cplx<T>* const* src2 = nullptr);
    void pad_inv_big(int njobs, const cplx<T>* const* src, int lev0, int nlev);
    void fwd_big_unpad(int njobs, const T* const* src, cplx<T>* const* dst, int lev0, int nlev);
    void cols_unpad(int njobs, cplx<T>* const* dst, int lev0, int nlev);

    // subsystems
    void filt_da(int f, int dfdx, int dfdy, cplx<T>* keep = nullptr);
    void ddx(int f, int dfdx);
    void ddy(int f, int dfdy);
    void ddxy(int f, int dfdx, int dfdy);
    void ddz_uv(int f, int dfdz);
    void ddz_w(int f, int dfdz);
    void ddz_uvw();        // ddz_uv(u), ddz_uv(v), ddz_w(w): one pass, the same values
    void test_filter_plane(const T* in, T* out, int op);
    void wallstress(const cplx<T>* su = nullptr, const cplx<T>* sv = nullptr);
    void calc_sij();
    void sgs_stag();
    void interpolag_sdep(bool in_place = true);
    void lagrange_sdep();
    void divstress_uv();
    void divstress_w();
    // the step's fused forms: filt_da of u, v, w in one pass pair; both
    // divstress routines with one forward and one inverse pass
    void filt_da_uvw(bool forward = true);
    void divstress_uvw();
    void convec(bool reuse_spectra = false, bool spectral_out = false);
    void press_stag_array();
    void project();
    void rhs_ab2();
    void rhs_gradp();
    bool refresh_flags(bool force_lag = false);
    const double* clk() const { return cfl_mode ? d_clock : nullptr; }
    void clock_begin();
    void clock_reset(int m);
    void clock_get(int m, double* out5);
    void clock_set(int m, double dt, double total_time, bool fresh);
    void advance_clocks();
    void step_body();
    void step_body_spec(bool cold);
    void swap_rhs();
    int count_cold() const;
    void copy_spec_member(int src, int dst);
    void run_body(bool cold);
    void after_body();
    void divstress_spec();
    void press_spec();
    void materialize(unsigned groups);
    void touch(bool modify);
    void high_part(T* const* f, T* const* out, int nf, int ka, int kb);
    void rhs_high(T* const* f, T* const* t, T* const* g, T* const* scratch);
    void step();
    void step_graph();
    void drop_graphs();
    void turbines_forcing();
    void pull_turbine_state();
    void rebuild_turbines();
    double rmsdiv(int m = 0);
    double energy(int m = 0);
    double max_cfl(int m = 0);
    double tau_wall_bot(int m = 0);
    void fill_level(int id, int k, double v);
    struct Fill {
        int id, k;
        double v;
    };
    // the planes (id, k) of every stepped member set to v, one launch
    void fill_levels(std::initializer_list<Fill> f);
    void copy_member(int src, int dst);
    void set_member_params(int m, double zo_m, double mpfx, double mpfy);
    void seed_member(int m, long seed, double noise);
    // Remove member `slot`: the last stepped member moves into it.
    void retire_member(int slot);
    void check_device_error();
    void stats_enable(int nstart, int nend, int nskip);
    void stats_after_step();
    double stats_time(int m);
    void stats_mean(int m, int q, double* host);
    void stats_sums(int m, int q, double* host);
    void stats_restore(int m, double total_time, const double* sums);
    void stats_spectrum(int m, int c, double* host);
};

// Sums of (a - b)^2 and b^2 of one field over i < nx, k0..k1 (two cases of
// one grid, any precisions).
template <typename TA, typename TB>
void diff_l2(Case<TA>& a, int ma, Case<TB>& b, int mb, int field, int k0, int k1, double* num,
             double* den);

// Device bytes a batch of `members` allocates (tsize = 8 or 4).
size_t case_bytes(const lg_params& p, int members, int tsize);

// Time every launch choice of c's passes and record the fastest in the
// tuning table; returns the number of passes tuned.
template <typename T>
int autotune(Case<T>& c, int reps);

// Replace member m's actuator disks (nturb = 0 removes them).
template <typename T>
void turbines_init(Case<T>& c, int m, int nturb, const double* spec8, double alpha1,
                   double alpha2, double filter_cutoff, int adm_correction);
template <typename T>
void turbines_set_ct(Case<T>& c, int m, const double* ct);
template <typename T>
void turbines_set_udt(Case<T>& c, int m, const double* udt);

}  // namespace lg
