// extern "C" boundary: see include/lesgo_amd.h for the per-function
// reference citations.
#include <cmath>
#include <cstdint>
#include <new>
#include <string>
#include <vector>

#include "case.hpp"

namespace lg {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace lg

struct lg_case {
    int precision = LG_FP64;
    lg::Case<double>* c64 = nullptr;
    lg::Case<float>* c32 = nullptr;
};

namespace {

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return LG_OK;
    } catch (const lg::ArgError&) {
        return LG_ERR_ARG;
    } catch (const lg::HipError&) {
        return LG_ERR_HIP;
    } catch (const lg::PivotError&) {
        return LG_ERR_PIVOT;
    } catch (const std::bad_alloc&) {
        lg::set_error("host allocation failed");
        return LG_ERR_ALLOC;
    } catch (...) {
        lg::set_error("unknown error");
        return LG_ERR_STATE;
    }
}

#define REQUIRE(cond, msg)          \
    do {                            \
        if (!(cond)) {              \
            lg::set_error(msg);     \
            throw lg::ArgError();   \
        }                           \
    } while (0)

// Run fn(Case<T>&) on whichever precision the handle holds.
template <typename F>
int dispatch(lg_case* h, F&& fn) {
    if (!h) {
        lg::set_error("null case handle");
        return LG_ERR_ARG;
    }
    return guarded([&] {
        if (h->c64) fn(*h->c64);
        else fn(*h->c32);
    });
}

// Entry points that read physical fields make the stale ones first; those
// that change them also end the spectral state (the next step starts from
// the physical fields).
template <typename F>
int dispatch_mod(lg_case* h, F&& fn) {
    return dispatch(h, [&](auto& c) {
        c.touch(true);
        fn(c);
    });
}

template <typename F>
int dispatch_read(lg_case* h, F&& fn) {
    return dispatch(h, [&](auto& c) {
        c.touch(false);
        fn(c);
    });
}

bool smooth(int n) {
    if (n <= 0) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}

void check_field(int id) { REQUIRE(id >= 0 && id < LG_NFIELDS, "field id out of range"); }

template <typename C>
void check_member(const C& c, int m) {
    REQUIRE(m >= 0 && m < c.M, "member index out of range");
}

}  // namespace

extern "C" {

const char* lg_last_error(void) { return lg::g_err.c_str(); }

int lg_create_batch(const lg_params* p, int members, lg_case** out) {
    return guarded([&] {
        REQUIRE(p && out, "null argument");
        REQUIRE(members >= 1, "members must be >= 1");
        REQUIRE((long)(p->Nz + 2) * members <= 65535,
                "members * (Nz + 2) must not exceed 65535 (launch grid limit)");
        REQUIRE(p->nx >= 4 && p->ny >= 4 && p->nx % 4 == 0 && p->ny % 4 == 0,
                "nx and ny must be multiples of 4 (3/2-rule grids must stay even)");
        REQUIRE(smooth(p->nx) && smooth(p->ny), "nx and ny must be of the form 2^a 3^b");
        REQUIRE(p->Nz >= 3, "Nz must be at least 3");
        REQUIRE(p->lx > 0 && p->ly > 0 && p->lz > 0 && p->dt > 0 && p->zo > 0,
                "lengths, dt and zo must be positive");
        REQUIRE(p->cs_count >= 1, "cs_count must be >= 1");
        REQUIRE(p->precision == LG_FP64 || p->precision == LG_MIXED, "bad precision");
        REQUIRE(!p->use_cfl_dt || p->cfl > 0, "cfl must be > 0 with use_cfl_dt");
        auto* h = new lg_case();
        h->precision = p->precision;
        try {
            if (p->precision == LG_FP64) h->c64 = new lg::Case<double>(*p, members);
            else h->c32 = new lg::Case<float>(*p, members);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int lg_create(const lg_params* p, lg_case** out) { return lg_create_batch(p, 1, out); }

int lg_destroy(lg_case* h) {
    if (!h) return LG_OK;
    int rc = guarded([&] {
        delete h->c64;
        delete h->c32;
    });
    delete h;
    return rc;
}

int lg_sizes(const lg_case* h, int* nz, int* ld, int* ny) {
    if (!h) return LG_ERR_ARG;
    const lg::Dims& d = h->c64 ? h->c64->d : h->c32->d;
    if (nz) *nz = d.nz;
    if (ld) *ld = d.ld;
    if (ny) *ny = d.ny;
    return LG_OK;
}

int lg_upload_member(lg_case* h, int m, int field, const double* host) {
    return dispatch_mod(h, [&](auto& c) {
        check_field(field);
        check_member(c, m);
        REQUIRE(host, "null host pointer");
        using T = typename std::remove_reference<decltype(*c.fld[0])>::type;
        if constexpr (std::is_same<T, double>::value) {
            LG_HIP(hipMemcpyAsync(c.mem(field, m), host, c.d.field * sizeof(double),
                                  hipMemcpyHostToDevice, c.stream));
        } else {
            std::vector<T> tmp((size_t)c.d.field);
            for (long i = 0; i < c.d.field; ++i) tmp[i] = (T)host[i];
            LG_HIP(hipMemcpyAsync(c.mem(field, m), tmp.data(), c.d.field * sizeof(T),
                                  hipMemcpyHostToDevice, c.stream));
        }
        LG_HIP(hipStreamSynchronize(c.stream));
    });
}

int lg_download_member(lg_case* h, int m, int field, double* host) {
    return dispatch_read(h, [&](auto& c) {
        check_field(field);
        check_member(c, m);
        REQUIRE(host, "null host pointer");
        using T = typename std::remove_reference<decltype(*c.fld[0])>::type;
        if constexpr (std::is_same<T, double>::value) {
            LG_HIP(hipMemcpyAsync(host, c.mem(field, m), c.d.field * sizeof(double),
                                  hipMemcpyDeviceToHost, c.stream));
            LG_HIP(hipStreamSynchronize(c.stream));
        } else {
            std::vector<T> tmp((size_t)c.d.field);
            LG_HIP(hipMemcpyAsync(tmp.data(), c.mem(field, m), c.d.field * sizeof(T),
                                  hipMemcpyDeviceToHost, c.stream));
            LG_HIP(hipStreamSynchronize(c.stream));
            for (long i = 0; i < c.d.field; ++i) host[i] = (double)tmp[i];
        }
    });
}

int lg_upload(lg_case* h, int field, const double* host) {
    return lg_upload_member(h, 0, field, host);
}

int lg_download(lg_case* h, int field, double* host) {
    return lg_download_member(h, 0, field, host);
}

int lg_seed_member(lg_case* h, int m, long seed, double noise) {
    return dispatch(h, [&](auto& c) {
        check_member(c, m);
        c.seed_member(m, seed, noise);
    });
}

int lg_seed_state(lg_case* h, long seed, double noise) {
    const int rc = lg_seed_member(h, 0, seed, noise);
    if (rc != LG_OK) return rc;
    return lg_set_counters(h, 0, 0);
}

int lg_set_counters(lg_case* h, int jt, int jt_total) {
    return dispatch(h, [&](auto& c) {
        c.jt = jt;
        c.jt_total = jt_total;
        for (int m = 0; m < c.M; ++m) c.mjt[m] = jt;
    });
}

int lg_advance_clock(lg_case* h) {
    return dispatch_mod(h, [&](auto& c) {
        c.refresh_flags();
        c.advance_clocks();
        c.clock_begin();
        LG_HIP(hipGetLastError());
    });
}

int lg_resync(lg_case* h) {
    return dispatch_mod(h, [&](auto& c) { (void)c; });
}

int lg_get_counters(const lg_case* h, int* jt, int* jt_total) {
    if (!h) return LG_ERR_ARG;
    if (h->c64) {
        if (jt) *jt = h->c64->jt;
        if (jt_total) *jt_total = h->c64->jt_total;
    } else {
        if (jt) *jt = h->c32->jt;
        if (jt_total) *jt_total = h->c32->jt_total;
    }
    return LG_OK;
}

int lg_synchronize(lg_case* h) {
    return dispatch(h, [&](auto& c) {
        LG_HIP(hipStreamSynchronize(c.stream));
        c.check_device_error();
    });
}

int lg_filt_da(lg_case* h, int f, int dfdx, int dfdy) {
    return dispatch_mod(h, [&](auto& c) {
        check_field(f);
        check_field(dfdx);
        check_field(dfdy);
        c.filt_da(f, dfdx, dfdy);
    });
}

int lg_ddx(lg_case* h, int f, int dfdx) {
    return dispatch_mod(h, [&](auto& c) {
        check_field(f);
        check_field(dfdx);
        c.ddx(f, dfdx);
    });
}

int lg_ddy(lg_case* h, int f, int dfdy) {
    return dispatch_mod(h, [&](auto& c) {
        check_field(f);
        check_field(dfdy);
        c.ddy(f, dfdy);
    });
}

int lg_ddxy(lg_case* h, int f, int dfdx, int dfdy) {
    return dispatch_mod(h, [&](auto& c) {
        check_field(f);
        check_field(dfdx);
        check_field(dfdy);
        c.ddxy(f, dfdx, dfdy);
    });
}

int lg_test_filter(lg_case* h, int width, double* plane) {
    return dispatch(h, [&](auto& c) {
        REQUIRE(width == 2 || width == 4, "test filter width must be 2 or 4 (delta)");
        REQUIRE(plane, "null host pointer");
        using T = typename std::remove_reference<decltype(*c.fld[0])>::type;
        const long n = c.d.plane;
        std::vector<T> buf((size_t)n);
        for (long i = 0; i < n; ++i) buf[i] = (T)plane[i];
        T* dp = nullptr;
        LG_HIP(hipMalloc(&dp, 2 * n * sizeof(T)));
        try {
            LG_HIP(hipMemcpyAsync(dp, buf.data(), n * sizeof(T), hipMemcpyHostToDevice,
                                  c.stream));
            c.test_filter_plane(dp, dp + n, width == 2 ? lg::OP_G1 : lg::OP_G2);
            LG_HIP(hipMemcpyAsync(buf.data(), dp + n, n * sizeof(T), hipMemcpyDeviceToHost,
                                  c.stream));
            LG_HIP(hipStreamSynchronize(c.stream));
        } catch (...) {
            (void)hipFree(dp);
            throw;
        }
        LG_HIP(hipFree(dp));
        for (long i = 0; i < n; ++i) plane[i] = (double)buf[i];
    });
}

int lg_ddz_uv(lg_case* h, int f, int dfdz) {
    return dispatch_mod(h, [&](auto& c) {
        check_field(f);
        check_field(dfdz);
        c.ddz_uv(f, dfdz);
    });
}

int lg_ddz_w(lg_case* h, int f, int dfdz) {
    return dispatch_mod(h, [&](auto& c) {
        check_field(f);
        check_field(dfdz);
        c.ddz_w(f, dfdz);
    });
}

int lg_wallstress(lg_case* h) { return dispatch_mod(h, [&](auto& c) { c.wallstress(); }); }
int lg_sgs_stag(lg_case* h) {
    return dispatch_mod(h, [&](auto& c) {
        c.refresh_flags();
        c.sgs_stag();
    });
}
int lg_lagrange_sdep(lg_case* h) {
    return dispatch_mod(h, [&](auto& c) {
        c.refresh_flags(true);
        c.lagrange_sdep();
    });
}
int lg_interpolag_sdep(lg_case* h) {
    return dispatch_mod(h, [&](auto& c) {
        c.refresh_flags(true);
        c.interpolag_sdep();
    });
}
int lg_divstress_uv(lg_case* h) { return dispatch_mod(h, [&](auto& c) { c.divstress_uv(); }); }
int lg_divstress_w(lg_case* h) { return dispatch_mod(h, [&](auto& c) { c.divstress_w(); }); }
int lg_convec(lg_case* h) { return dispatch_mod(h, [&](auto& c) { c.convec(); }); }

int lg_press_stag_array(lg_case* h) {
    return dispatch_mod(h, [&](auto& c) {
        c.press_stag_array();
        c.check_device_error();
    });
}

int lg_project(lg_case* h) { return dispatch_mod(h, [&](auto& c) { c.project(); }); }

int lg_copy_field(lg_case* h, int src, int dst) {
    return dispatch_mod(h, [&](auto& c) {
        check_field(src);
        check_field(dst);
        using T = typename std::remove_reference<decltype(*c.fld[0])>::type;
        if (src != dst)
            LG_HIP(hipMemcpyAsync(c.fld[dst], c.fld[src], c.Ma * c.d.field * sizeof(T),
                                  hipMemcpyDeviceToDevice, c.stream));
    });
}

int lg_rhs_ab2(lg_case* h) {
    return dispatch_mod(h, [&](auto& c) {
        c.refresh_flags();
        c.rhs_ab2();
    });
}
int lg_rhs_gradp(lg_case* h) { return dispatch_mod(h, [&](auto& c) { c.rhs_gradp(); }); }

int lg_step(lg_case* h, int nsteps) {
    return dispatch(h, [&](auto& c) {
        REQUIRE(nsteps >= 0, "nsteps must be >= 0");
        for (int n = 0; n < nsteps; ++n) c.step();
        LG_HIP(hipGetLastError());
    });
}

#define LG_DIAG(NAME, FN)                                                  \
    int NAME##_member(lg_case* h, int m, double* out) {                    \
        return dispatch(h, [&](auto& c) {                                  \
            REQUIRE(out, "null output");                                   \
            check_member(c, m);                                            \
            *out = c.FN(m);                                                \
        });                                                                \
    }                                                                      \
    int NAME(lg_case* h, double* out) { return NAME##_member(h, 0, out); }
LG_DIAG(lg_rmsdiv, rmsdiv)
LG_DIAG(lg_energy, energy)
LG_DIAG(lg_max_cfl, max_cfl)
LG_DIAG(lg_tau_wall_bot, tau_wall_bot)
#undef LG_DIAG

int lg_clock_member(lg_case* h, int m, double* out) {
    return dispatch(h, [&](auto& c) {
        REQUIRE(out, "null output");
        check_member(c, m);
        c.clock_get(m, out);
    });
}

int lg_set_clock_member(lg_case* h, int m, double dt, double total_time, int fresh) {
    return dispatch(h, [&](auto& c) {
        check_member(c, m);
        REQUIRE(dt > 0 && std::isfinite(total_time), "dt must be > 0 and total_time finite");
        c.clock_set(m, dt, total_time, fresh != 0);
    });
}

int lg_turbines_init_member(lg_case* h, int m, int nturb, const double* spec8, double alpha1,
                            double alpha2, double filter_cutoff, int adm_correction) {
    return dispatch_read(h, [&](auto& c) {
        check_member(c, m);
        REQUIRE(nturb >= 0 && (nturb == 0 || spec8), "bad turbine list");
        REQUIRE(alpha1 > 0 && alpha2 > 0 && filter_cutoff >= 0, "bad turbine filter parameters");
        for (int s = 0; s < nturb; ++s) REQUIRE(spec8[8 * s + 3] > 0, "turbine diameter must be > 0");
        lg::turbines_init(c, m, nturb, spec8, alpha1, alpha2, filter_cutoff, adm_correction);
    });
}

int lg_turbines_init(lg_case* h, int nturb, const double* spec8, double alpha1, double alpha2,
                     double filter_cutoff, int adm_correction) {
    return lg_turbines_init_member(h, 0, nturb, spec8, alpha1, alpha2, filter_cutoff,
                                   adm_correction);
}

int lg_turbines_forcing(lg_case* h) {
    return dispatch_mod(h, [&](auto& c) {
        REQUIRE(c.turb, "turbines not initialised");
        // fxa, fya, fza are written on the disk nodes only: a buffer that
        // still holds forcing from before a disk-set change is cleared first
        if (c.fclear_pending > 0) {
            for (int id : {LG_FXA, LG_FYA, LG_FZA})
                LG_HIP(hipMemsetAsync(c.fld[id], 0, c.M * c.d.field * sizeof(*c.fld[id]),
                                      c.stream));
            c.fclear_pending = 0;
        }
        c.turbines_forcing();
    });
}

int lg_turbines_get_member(lg_case* h, int m, double* u_d_T, double* f_n, int nturb) {
    return dispatch(h, [&](auto& c) {
        check_member(c, m);
        REQUIRE(c.turb && nturb == c.mturb[m].n(), "turbine count mismatch");
        c.pull_turbine_state();
        const auto& st = c.mturb[m].state;
        for (int s = 0; s < nturb; ++s) {
            if (u_d_T) u_d_T[s] = st[3 * s + 1];
            if (f_n) f_n[s] = st[3 * s + 2];
        }
    });
}

int lg_turbines_history_member(lg_case* h, int m, double* out, int nturb) {
    return dispatch(h, [&](auto& c) {
        check_member(c, m);
        REQUIRE(out, "null output");
        REQUIRE(c.turb && c.turb->n > 0 && nturb == c.mturb[m].n(), "turbine count mismatch");
        c.pull_turbine_state();
        std::vector<double> ctr(3 * (size_t)nturb);
        LG_HIP(hipMemcpy(ctr.data(), c.turb->center + 3L * c.turb->first[m],
                         ctr.size() * sizeof(double), hipMemcpyDeviceToHost));
        const auto& st = c.mturb[m].state;
        for (int s = 0; s < nturb; ++s) {
            out[5 * s + 0] = ctr[3 * s + 0];
            out[5 * s + 1] = ctr[3 * s + 1];
            out[5 * s + 2] = ctr[3 * s + 2];
            out[5 * s + 3] = st[3 * s + 0];
            out[5 * s + 4] = st[3 * s + 1];
        }
    });
}

int lg_turbines_get(lg_case* h, double* u_d_T, double* f_n, int nturb) {
    return lg_turbines_get_member(h, 0, u_d_T, f_n, nturb);
}

int lg_turbines_time_filter(lg_case* h, double T_avg_dim, double z_i, double u_star) {
    return dispatch(h, [&](auto& c) {
        REQUIRE(z_i > 0 && u_star > 0, "z_i and u_star must be > 0");
        c.t_avg_dim = T_avg_dim;
        c.z_i = z_i;
        c.u_star = u_star;
        c.drop_graphs();
    });
}

int lg_turbines_set_ct_member(lg_case* h, int m, const double* ct_prime, int nturb) {
    return dispatch(h, [&](auto& c) {
        check_member(c, m);
        REQUIRE(ct_prime && nturb == c.mturb[m].n() && nturb > 0, "turbine count mismatch");
        lg::turbines_set_ct(c, m, ct_prime);
    });
}

int lg_turbines_set_udt_member(lg_case* h, int m, const double* u_d_T, int nturb) {
    return dispatch(h, [&](auto& c) {
        check_member(c, m);
        REQUIRE(u_d_T && nturb == c.mturb[m].n() && nturb > 0, "turbine count mismatch");
        lg::turbines_set_udt(c, m, u_d_T);
    });
}

int lg_batch_info(const lg_case* h, int* members, int* active) {
    if (!h) return LG_ERR_ARG;
    const int M = h->c64 ? h->c64->M : h->c32->M;
    const int Ma = h->c64 ? h->c64->Ma : h->c32->Ma;
    if (members) *members = M;
    if (active) *active = Ma;
    return LG_OK;
}

int lg_set_active(lg_case* h, int active) {
    return dispatch_read(h, [&](auto& c) {
        REQUIRE(active >= 1 && active <= c.M, "active member count out of range");
        c.Ma = active;
    });
}

int lg_member_params(lg_case* h, int m, double zo, double mean_p_force_x,
                     double mean_p_force_y) {
    return dispatch(h, [&](auto& c) {
        check_member(c, m);
        REQUIRE(zo > 0, "zo must be positive");
        c.set_member_params(m, zo, mean_p_force_x, mean_p_force_y);
    });
}

int lg_copy_member(lg_case* h, int src, int dst) {
    return dispatch(h, [&](auto& c) {
        check_member(c, src);
        check_member(c, dst);
        c.copy_member(src, dst);
        LG_HIP(hipStreamSynchronize(c.stream));
    });
}

int lg_autotune(lg_case* h, int reps, int* npasses) {
    return dispatch(h, [&](auto& c) {
        REQUIRE(reps >= 1, "reps must be >= 1");
        const int n = lg::autotune(c, reps);
        if (npasses) *npasses = n;
    });
}

int lg_tune_get(int kind, int precision, int N, int other, int* n1, int* n2, int* per_wg,
                int* variant) {
    return guarded([&] {
        REQUIRE(kind >= 0 && kind < lg::PK_NKINDS, "pass kind out of range");
        REQUIRE(precision == LG_FP64 || precision == LG_MIXED, "bad precision");
        const int ts = precision == LG_FP64 ? 8 : 4;
        lg::PassChoice c;
        if (!lg::tuned_choice(kind, ts, N, other, c))
            c = ts == 8 ? lg::default_choice<double>(kind, N, other)
                        : lg::default_choice<float>(kind, N, other);
        if (n1) *n1 = c.n1;
        if (n2) *n2 = c.n2;
        if (per_wg) *per_wg = c.S;
        if (variant) *variant = c.v;
    });
}

int lg_tune_set(int kind, int precision, int N, int other, int n1, int n2, int per_wg,
                int variant) {
    return guarded([&] {
        REQUIRE(kind >= 0 && kind < lg::PK_NKINDS, "pass kind out of range");
        REQUIRE(precision == LG_FP64 || precision == LG_MIXED, "bad precision");
        const lg::PassChoice c{n1, n2, per_wg, variant};
        const bool ok = precision == LG_FP64 ? lg::choice_ok<double>(kind, c, N, other)
                                             : lg::choice_ok<float>(kind, c, N, other);
        REQUIRE(ok, "launch choice not launchable for this pass");
        lg::set_tuned_choice(kind, precision == LG_FP64 ? 8 : 4, N, other, c);
    });
}

int lg_tune_clear(void) {
    return guarded([&] { lg::clear_tuned_choices(); });
}

int lg_tune_count(int* n) {
    return guarded([&] {
        REQUIRE(n, "null output");
        *n = lg::tuned_count();
    });
}

int lg_tune_entry(int i, int* kind, int* precision, int* N, int* other, int* n1, int* n2,
                  int* per_wg, int* variant) {
    return guarded([&] {
        int key[4];
        lg::PassChoice c;
        REQUIRE(lg::tuned_entry(i, key, c), "tuning entry out of range");
        if (kind) *kind = key[0];
        if (precision) *precision = key[1] == 8 ? LG_FP64 : LG_MIXED;
        if (N) *N = key[2];
        if (other) *other = key[3];
        if (n1) *n1 = c.n1;
        if (n2) *n2 = c.n2;
        if (per_wg) *per_wg = c.S;
        if (variant) *variant = c.v;
    });
}

int lg_stats_enable(lg_case* h, int nstart, int nend, int nskip) {
    return dispatch(h, [&](auto& c) {
        REQUIRE(nskip >= 1 && nend >= nstart, "bad statistics window");
        c.stats_enable(nstart, nend, nskip);
    });
}

int lg_stats_info(lg_case* h, double* total_time, int* initialized) {
    return dispatch(h, [&](auto& c) {
        if (total_time) *total_time = c.stats_time(0);
        if (initialized) *initialized = c.st.initialized ? 1 : 0;
    });
}

int lg_stats_time(lg_case* h, int m, double* total_time) {
    return dispatch(h, [&](auto& c) {
        check_member(c, m);
        REQUIRE(total_time, "null output pointer");
        *total_time = c.stats_time(m);
    });
}

int lg_stats_mean(lg_case* h, int m, int quantity, double* host) {
    return dispatch(h, [&](auto& c) {
        check_member(c, m);
        REQUIRE(host, "null host pointer");
        c.stats_mean(m, quantity, host);
    });
}

int lg_stats_sums(lg_case* h, int m, int quantity, double* host) {
    return dispatch(h, [&](auto& c) {
        check_member(c, m);
        REQUIRE(host, "null host pointer");
        c.stats_sums(m, quantity, host);
    });
}

int lg_stats_restore(lg_case* h, int m, double total_time, const double* sums) {
    return dispatch(h, [&](auto& c) {
        check_member(c, m);
        REQUIRE(sums, "null host pointer");
        REQUIRE(total_time >= 0.0, "averaging time must be >= 0");
        c.stats_restore(m, total_time, sums);
    });
}

int lg_stats_spectrum(lg_case* h, int m, int component, double* host) {
    return dispatch(h, [&](auto& c) {
        check_member(c, m);
        REQUIRE(host, "null host pointer");
        c.stats_spectrum(m, component, host);
    });
}

int lg_diff_l2(lg_case* a, int ma, lg_case* b, int mb, int field, int k0, int k1, double* num,
               double* den) {
    if (!a || !b) {
        lg::set_error("null case handle");
        return LG_ERR_ARG;
    }
    return guarded([&] {
        check_field(field);
        REQUIRE(num && den, "null output");
        auto run = [&](auto& ca, auto& cb) {
            check_member(ca, ma);
            check_member(cb, mb);
            REQUIRE(k0 >= 0 && k1 <= ca.d.nz && k0 <= k1, "level range out of bounds");
            ca.touch(false);
            cb.touch(false);
            lg::diff_l2(ca, ma, cb, mb, field, k0, k1, num, den);
        };
        if (a->c64 && b->c64) run(*a->c64, *b->c64);
        else if (a->c64) run(*a->c64, *b->c32);
        else if (b->c64) run(*a->c32, *b->c64);
        else run(*a->c32, *b->c32);
    });
}

int lg_step_graph(lg_case* h, int nsteps) {
    return dispatch(h, [&](auto& c) {
        REQUIRE(nsteps >= 0, "nsteps must be >= 0");
        for (int n = 0; n < nsteps; ++n) c.step_graph();
        LG_HIP(hipGetLastError());
    });
}

int lg_timing_enable(lg_case* h, int on) {
    return dispatch(h, [&](auto& c) {
        LG_HIP(hipStreamSynchronize(c.stream));
        c.timer.reset();
        c.timer.on = on != 0;
    });
}

int lg_timing_report(lg_case* h, int kind, double* ms, double* launches, double* bytes) {
    return dispatch(h, [&](auto& c) {
        REQUIRE(kind >= 0 && kind < lg::KK_NKINDS, "timing kind out of range");
        double a, b, e;
        c.timer.report(kind, a, b, e);
        if (ms) *ms = a;
        if (launches) *launches = b;
        if (bytes) *bytes = e;
    });
}

}  // extern "C"
