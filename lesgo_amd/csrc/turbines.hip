// Actuator-disk model (turbines.f90:123-638, turbine_indicator.f90:43-169).
// Initialisation (indicator table, node lists) runs once on the host; the
// per-step disk average, thrust and force scatter run on the device.
#include <algorithm>
#include <cmath>
#include <vector>

#include "case.hpp"
#include "pointwise.hpp"

namespace lg {

static constexpr double PI_T = 3.1415926535897932384626433;

// turb_ind_func_t (turbine_indicator.f90:43-169).  R2 is row 0 of the
// circular convolution of the Gaussian g (width delta2) with the unit-disk
// indicator h on an N x N grid.  The Gaussian factorises, so the row is
// evaluated as sum_l e_l H1[(j - l) mod N] with H1[m] = sum_k e_k
// h[(-k) mod N][m]: exact in real arithmetic, O(N^2).
struct Indicator {
    double ell, d1, d2, thk, s6d1, M;
    std::vector<double> r, R2;

    Indicator(double delta1, double delta2, double thk_in, double dia, int N = 2048) {
        ell = 0.5 * dia;
        d1 = delta1 / ell;
        d2 = delta2 / ell;
        thk = thk_in / ell;
        s6d1 = std::sqrt(6.0) / d1;
        const double L = 1 + 10 * d2 / std::sqrt(12.0);
        const double d = 2 * L / N;
        std::vector<double> yz(N), e(N);
        for (int i = 0; i < N; ++i) {
            yz[i] = -L + d * ((i + 1) - 0.5);
            e[i] = std::exp(-6 * (yz[i] * yz[i]) / (d2 * d2));
        }
        const double cg = 6.0 / (PI_T * d2 * d2);
        std::vector<double> H1(N, 0.0);
        for (int k = 0; k < N; ++k) {
            const int a = (N - k) % N;
            const double ya2 = yz[a] * yz[a];
            double* out = H1.data();
            for (int m = 0; m < N; ++m) {
                if (std::sqrt(ya2 + yz[m] * yz[m]) < 1) out[m] += e[k] * (1.0 / PI_T);
            }
        }
        r.resize(N / 2);
        R2.resize(N / 2);
        for (int j = 0; j < N / 2; ++j) {
            double s = 0.0;
            for (int l = 0; l < N; ++l) s += e[l] * H1[(j - l + N) % N];
            R2[j] = d * d * cg * s;
            r[j] = yz[j] + L;
        }
        const double dr = r[1] - r[0];
        double norm = 0.0;
        for (int j = 0; j < N / 2; ++j) norm += 2 * PI_T * r[j] * R2[j] * dr;
        for (auto& v : R2) v /= norm;
        double m = 0.0;
        for (int j = 0; j < N / 2; ++j) m += 2 * PI_T * r[j] * R2[j] * R2[j] * dr;
        M = m * PI_T;
    }

    // linear_interp_sa + binary_search (functions.f90:711-860)
    double interp(double xq) const {
        const int N = (int)r.size();
        if (xq < r[0]) return R2[0];
        if (xq > r[N - 1]) return R2[N - 1];
        int low = 1, high = N;          // 1-based, as binary_search
        while (high - low > 1) {
            const int mid = (low + high) / 2;
            if (r[mid - 1] > xq) high = mid;
            else if (r[mid - 1] < xq) low = mid;
            else {
                low = mid;
                break;
            }
        }
        if (low == N) return R2[N - 1];
        const int i = low - 1;
        return R2[i] + (xq - r[i]) * (R2[i + 1] - R2[i]) / (r[i + 1] - r[i]);
    }

    double val(double rr_in, double x_in) const {
        const double rr = rr_in / ell;
        const double xx = x_in / ell;
        const double r2 = interp(rr);
        const double R1 = 0.5 / thk * (std::erf(s6d1 * (xx + 0.5 * thk)) -
                                       std::erf(s6d1 * (xx - 0.5 * thk)));
        return R1 * r2 / (ell * ell * ell);
    }
};

static int fnint(double x) {
    return (int)std::floor(std::fabs(x) + 0.5) * (x >= 0 ? 1 : -1);
}

// Disk-averaged velocity, thrust (turbines.f90:517-580): one block per
// turbine.  The u_d_T filter weight eps = (dt_dim/T)/(1 + dt_dim/T),
// dt_dim = dt z_i / u_star, for T = T_avg_dim > 0, else 1 (561-566); dt is
// the member's clock with a CFL-adaptive dt.
struct TurbFilter {
    double T_avg_dim, z_i, u_star, dt;
    const double* clk;
};

// The disk centre samples written with the history records (turbines.f90:
// 538-547): u, v and w_uv at each disk's centre cell.
struct TurbCenter {
    const long* lin;                 // u, v offset of the centre cell
    const long* wlin;                // w offset of the level below its w_uv
    double* out;                     // 3 per turbine
};

template <typename T>
__global__ void k_turb_reduce(const T* u, const T* v, const T* w, const long* node_lin,
                              const double* node_ind, const int* off, const double* coef,
                              double* state, long plane, double dV, TurbFilter tf,
                              TurbCenter tc) {
    const int s = blockIdx.x;
    const double n1 = coef[TC_N * s + 0], n2 = coef[TC_N * s + 1], n3 = coef[TC_N * s + 2];
    double acc = 0.0;
    for (int l = off[s] + threadIdx.x; l < off[s + 1]; l += blockDim.x) {
        const long o = node_lin[l];
        const double wuv = 0.5 * ((double)w[o + plane] + (double)w[o]);
        acc += dV * node_ind[l] * (n1 * (double)u[o] + n2 * (double)v[o] + n3 * wuv);
    }
    // the disk's sum by wave shuffles, then across the block's waves
    acc = block_reduce(acc, RedSum{}, 0.0);
    if (threadIdx.x == 0) {
        const double ct = coef[TC_N * s + 3], dia = coef[TC_N * s + 4],
                     corr = coef[TC_N * s + 5];
        double ud = acc;
        ud = ud / corr;
        double eps = 1.0;
        if (tf.T_avg_dim > 0.0) {
            const int mem = (int)coef[TC_N * s + 6];
            const double dt = tf.clk ? tf.clk[(long)CK_N * mem + CK_DT] : tf.dt;
            const double dt_dim = dt * tf.z_i / tf.u_star;
            eps = (dt_dim / tf.T_avg_dim) / (1. + dt_dim / tf.T_avg_dim);
        }
        const double udt = (1. - eps) * state[3 * s + 1] + eps * ud;
        const double fn = -0.5 * ct * fabs(udt) * udt * 0.25 * 3.1415926535897932384626433 * (dia * dia);
        state[3 * s + 0] = ud;
        state[3 * s + 1] = udt;
        state[3 * s + 2] = fn;
        const long c = tc.lin[s], cw = tc.wlin[s];
        tc.out[3 * s + 0] = (double)u[c];
        tc.out[3 * s + 1] = (double)v[c];
        tc.out[3 * s + 2] = 0.5 * ((double)w[cw + plane] + (double)w[cw]);
    }
}

// Scatter f_n n ind to the nodes (turbines.f90:597-617).
template <typename T>
__global__ void k_turb_scatter(T* fx, T* fy, T* fzu, const long* node_lin, const double* node_ind,
                               const int* node_tid, const double* coef, const double* state,
                               int total) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= total) return;
    const int s = node_tid[l];
    const double fn = state[3 * s + 2];
    const long o = node_lin[l];
    fx[o] = (T)(fn * coef[TC_N * s + 0] * node_ind[l]);
    fy[o] = (T)(fn * coef[TC_N * s + 1] * node_ind[l]);
    fzu[o] = (T)(fn * coef[TC_N * s + 2] * node_ind[l]);
}

// fza = interp_to_w_grid(fza) (functions.f90:131-185), k = 1..nz.
template <typename T>
__global__ void k_interp_w(const T* src, T* dst, Geo g) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.ld, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    dst[o] = (T)0.5 * (src[o - g.plane] + src[o]);
}

template <typename T>
void turbines_init(Case<T>& c, int m, int nturb, const double* spec8, double alpha1,
                   double alpha2, double filter_cutoff, int adm_correction) {
    const Dims& d = c.d;
    MemberTurbines mt;
    mt.spec.assign(spec8, spec8 + 8 * nturb);
    mt.off.push_back(0);
    const double h = std::sqrt(d.dx * d.dx + d.dy * d.dy + d.dz * d.dz);
    const double delta1 = alpha1 * h, delta2 = alpha2 * h;
    const int nx = d.nx, ny = d.ny;
    for (int s = 0; s < nturb; ++s) {
        const double* t = spec8 + 8 * s;
        const double xloc = t[0], yloc = t[1], height = t[2], dia = t[3];
        const double thk = std::max(t[4], d.dx * 1.01);
        const double th1 = t[5], th2 = t[6], ctp = t[7];
        Indicator itf(delta1, delta2, thk, dia);
        mt.M.push_back(itf.M);
        const double search_rad = 0.5 * dia + 3 * std::max(alpha1, alpha2) * h;
        const int imax = std::min((int)(search_rad / d.dx + 2), nx / 2);
        const int jmax = std::min((int)(search_rad / d.dy + 2), ny / 2);
        const double n1 = -std::cos(PI_T * th1 / 180.) * std::cos(PI_T * th2 / 180.);
        const double n2 = -std::sin(PI_T * th1 / 180.) * std::cos(PI_T * th2 / 180.);
        const double n3 = std::sin(PI_T * th2 / 180.);
        const int icp = fnint(xloc / d.dx);
        const int jcp = fnint(yloc / d.dy);
        const double filt_max = itf.val(0.0, 0.0);
        std::vector<double> filt_l;
        std::vector<long> lin_l;
        double sumA = 0.0;
        for (int k = 1; k <= d.nz - 1; ++k) {
            for (int j = jcp - jmax + 1; j <= jcp + jmax; ++j) {
                for (int i = icp - imax + 1; i <= icp + imax; ++i) {
                    int i2, j2;
                    double rx, ry;
                    if (i < 1) {
                        i2 = ((i + nx - 1) % nx + nx) % nx + 1;
                        rx = ((i2 - 1) * d.dx - d.prm.lx) - xloc;
                    } else if (i > nx) {
                        i2 = (i + nx - 1) % nx + 1;
                        rx = (d.prm.lx + (i2 - 1) * d.dx) - xloc;
                    } else {
                        i2 = i;
                        rx = (i - 1) * d.dx - xloc;
                    }
                    if (j < 1) {
                        j2 = ((j + ny - 1) % ny + ny) % ny + 1;
                        ry = ((j2 - 1) * d.dy - d.prm.ly) - yloc;
                    } else if (j > ny) {
                        j2 = (j + ny - 1) % ny + 1;
                        ry = (d.prm.ly + (j2 - 1) * d.dy) - yloc;
                    } else {
                        j2 = j;
                        ry = (j - 1) * d.dy - yloc;
                    }
                    const double rz = (k - 0.5) * d.dz - height;
                    const double r = std::sqrt(rx * rx + ry * ry + rz * rz);
                    const double r_norm = std::fabs(rx * n1 + ry * n2 + rz * n3);
                    const double r_disk = std::sqrt(r * r - r_norm * r_norm);
                    const double f = itf.val(r_disk, r_norm);
                    if (f > filter_cutoff * filt_max) {
                        filt_l.push_back(f);
                        lin_l.push_back(((long)k * ny + (j2 - 1)) * d.ld + (i2 - 1));
                        sumA = sumA + f * d.dx * d.dy * d.dz;
                    }
                }
            }
        }
        for (size_t l = 0; l < filt_l.size(); ++l) {
            mt.ind.push_back(filt_l[l] / sumA);
            mt.lin.push_back(lin_l[l]);
        }
        mt.off.push_back((int)mt.lin.size());
        const double corr = adm_correction ? (1 + 0.25 * (1 - itf.M) * ctp) : 1.0;
        mt.coef.insert(mt.coef.end(), {n1, n2, n3, ctp, dia, corr, (double)m, 0.0});
        // u_d_T without a u_d_T.dat (turbines.f90:245-250)
        mt.state.insert(mt.state.end(), {0.0, -1.0, 0.0});
    }
    mt.adm = adm_correction != 0;
    c.pull_turbine_state();     // the other members keep their disk state
    c.mturb[m] = std::move(mt);
    c.rebuild_turbines();
}

// Copy the device disk state (u_d, u_d_T, f_n) into the members' host lists.
template <typename T>
void Case<T>::pull_turbine_state() {
    LG_HIP(hipStreamSynchronize(stream));
    if (!turb || turb->n == 0) return;
    std::vector<double> st(3 * (size_t)turb->n);
    LG_HIP(hipMemcpy(st.data(), turb->state, st.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int m = 0; m < M; ++m) {
        const int f0 = turb->first[m];
        std::copy(st.begin() + 3 * f0, st.begin() + 3 * turb->first[m + 1],
                  mturb[m].state.begin());
    }
}

// Concatenate every member's disks into the device node lists.  The forcing
// fields are cleared: they are zero off the nodes and rewritten on the nodes
// every step.
template <typename T>
void Case<T>::rebuild_turbines() {
    // the physical fields are made before the forcing arrays are cleared
    touch(false);
    LG_HIP(hipStreamSynchronize(stream));
    drop_graphs();
    turb.reset();
    auto ts = std::make_unique<TurbineSet>();
    std::vector<long> lin;
    std::vector<int> tid, off(1, 0);
    std::vector<double> ind, coef, st;
    std::vector<long> clin, cwlin;
    ts->first.push_back(0);
    int s = 0;
    int kmin = d.nz + 1, kmax = 0;
    for (int m = 0; m < M; ++m) {
        const MemberTurbines& mt = mturb[m];
        for (int q = 0; q < mt.n(); ++q, ++s) {
            for (int l = mt.off[q]; l < mt.off[q + 1]; ++l) {
                const int k = (int)(mt.lin[l] / d.plane);
                kmin = std::min(kmin, k);
                kmax = std::max(kmax, k);
                lin.push_back((long)m * d.field + mt.lin[l]);
                ind.push_back(mt.ind[l]);
                tid.push_back(s);
            }
            off.push_back((int)lin.size());
            coef.insert(coef.end(), mt.coef.begin() + TC_N * q, mt.coef.begin() + TC_N * (q + 1));
            coef[coef.size() - TC_N + 6] = (double)m;
            st.insert(st.end(), mt.state.begin() + 3 * q, mt.state.begin() + 3 * q + 3);
            // centre cell (turbines.f90:196-198: Fortran nint of xloc/dx, yloc/dy,
            // height/dz + 0.5); a centre index past the grid (0 or nx + 1) is
            // taken periodically and the level kept within 1..nz
            const double* sp = &mt.spec[8 * q];
            const long ic = ((std::lround(sp[0] / d.dx) - 1) % d.nx + d.nx) % d.nx;
            const long jc = ((std::lround(sp[1] / d.dy) - 1) % d.ny + d.ny) % d.ny;
            const long kc = std::min<long>(std::max<long>(std::lround(sp[2] / d.dz + 0.5), 1), d.nz);
            const long kw = std::min<long>(kc, d.nz - 1);
            clin.push_back((long)m * d.field + (kc * d.ny + jc) * d.ld + ic);
            cwlin.push_back((long)m * d.field + (kw * d.ny + jc) * d.ld + ic);
            if (mt.coef[TC_N * q + 2] != 0.0) ts->has_fz = true;
        }
        ts->first.push_back(s);
    }
    // The forcing is written on the nodes only, so the buffer the next step
    // writes is cleared here.  In the spectral step that is fprev (the
    // buffers exchange at the step's start); fxa/fya/fza keep the last
    // step's forcing, whose high part the next step's physical fields
    // still carry, and are cleared when they next become the write target.
    if (spec_on) {
        for (int q = 0; q < 3; ++q)
            LG_HIP(hipMemsetAsync(fprev[q], 0, M * d.field * sizeof(T), stream));
        fclear_pending = 2;
    } else {
        for (int id : {LG_FXA, LG_FYA, LG_FZA})
            LG_HIP(hipMemsetAsync(fld[id], 0, M * d.field * sizeof(T), stream));
    }
    LG_HIP(hipStreamSynchronize(stream));
    // levels where the forcing can be non-zero: the nodes' levels and, for
    // fza on the w grid, one above
    kf0 = 1;
    kf1 = 0;
    if (kmax >= kmin) {
        kf0 = std::max(1, kmin);
        kf1 = std::min(d.nz, kmax + 1);
    }
    if (s == 0) return;
    ts->n = s;
    ts->total_nodes = (int)lin.size();
    auto up = [&](auto*& dptr, const auto& hv) {
        using E = typename std::decay<decltype(hv[0])>::type;
        LG_HIP(hipMalloc(&dptr, std::max<size_t>(1, hv.size()) * sizeof(E)));
        if (!hv.empty())
            LG_HIP(hipMemcpy(dptr, hv.data(), hv.size() * sizeof(E), hipMemcpyHostToDevice));
    };
    up(ts->node_lin, lin);
    up(ts->node_ind, ind);
    up(ts->node_tid, tid);
    up(ts->off, off);
    up(ts->coef, coef);
    up(ts->state, st);
    up(ts->center_lin, clin);
    up(ts->center_wlin, cwlin);
    up(ts->center, std::vector<double>(3 * (size_t)s, 0.0));
    if (ts->has_fz) {
        LG_HIP(hipMalloc(&ts->fz_uv, M * d.field * sizeof(T)));
        LG_HIP(hipMemsetAsync(ts->fz_uv, 0, M * d.field * sizeof(T), stream));
        LG_HIP(hipStreamSynchronize(stream));
    }
    turb = std::move(ts);
}

template <typename T>
void Case<T>::turbines_forcing() {
    if (!turb) return;
    TurbineSet& t = *turb;
    if (t.total_nodes == 0) return;
    hipLaunchKernelGGL(k_turb_reduce<T>, dim3(t.n), dim3(256), 0, stream, (const T*)fld[LG_U],
                       (const T*)fld[LG_V], (const T*)fld[LG_W], (const long*)t.node_lin,
                       (const double*)t.node_ind, (const int*)t.off, (const double*)t.coef,
                       t.state, d.plane, d.dx * d.dy * d.dz,
                       TurbFilter{t_avg_dim, z_i, u_star, d.prm.dt, clk()},
                       TurbCenter{t.center_lin, t.center_wlin, t.center});
    // the node set is static, so fxa/fya/fza are zero off the nodes
    T* fz_uv = t.has_fz ? static_cast<T*>(t.fz_uv) : fld[LG_FZA];
    hipLaunchKernelGGL(k_turb_scatter<T>, dim3((t.total_nodes + 255) / 256), dim3(256), 0, stream,
                       fld[LG_FXA], fld[LG_FYA], fz_uv, (const long*)t.node_lin,
                       (const double*)t.node_ind, (const int*)t.node_tid, (const double*)t.coef,
                       (const double*)t.state, t.total_nodes);
    if (t.has_fz) {
        Geo g{d.nx, d.ny, d.nz, d.ld, d.plane, 1, d.nz, d.field};
        hipLaunchKernelGGL(k_interp_w<T>, dim3((unsigned)((d.ld * d.ny + 255) / 256), d.nz * Ma),
                           dim3(256), 0, stream, (const T*)fz_uv, fld[LG_FZA], g);
    }
}

// Dynamic Ct' (turbines.f90:504-515): new thrust coefficients of member m's
// disks in place (and the ADM correction that depends on them); the node
// lists stay.
template <typename T>
void turbines_set_ct(Case<T>& c, int m, const double* ct) {
    MemberTurbines& mt = c.mturb[m];
    const int n = mt.n();
    if (n == 0) return;
    for (int q = 0; q < n; ++q) {
        double* cf = &mt.coef[TC_N * q];
        cf[3] = ct[q];
        cf[5] = mt.adm ? (1 + 0.25 * (1 - mt.M[q]) * ct[q]) : 1.0;
        mt.spec[8 * q + 7] = ct[q];
    }
    LG_HIP(hipStreamSynchronize(c.stream));
    LG_HIP(hipMemcpy(c.turb->coef + (long)TC_N * c.turb->first[m], mt.coef.data(),
                     mt.coef.size() * sizeof(double), hipMemcpyHostToDevice));
}

// u_d_T of member m's disks (turbine_vel_init from u_d_T.dat, turbines.f90:
// 229-250).
template <typename T>
void turbines_set_udt(Case<T>& c, int m, const double* udt) {
    const int n = c.mturb[m].n();
    if (n == 0) return;
    c.pull_turbine_state();
    for (int q = 0; q < n; ++q) c.mturb[m].state[3 * q + 1] = udt[q];
    LG_HIP(hipMemcpy(c.turb->state + 3L * c.turb->first[m], c.mturb[m].state.data(),
                     3 * n * sizeof(double), hipMemcpyHostToDevice));
}

template void turbines_set_ct<double>(Case<double>&, int, const double*);
template void turbines_set_ct<float>(Case<float>&, int, const double*);
template void turbines_set_udt<double>(Case<double>&, int, const double*);
template void turbines_set_udt<float>(Case<float>&, int, const double*);
template void turbines_init<double>(Case<double>&, int, int, const double*, double, double,
                                    double, int);
template void turbines_init<float>(Case<float>&, int, int, const double*, double, double, double,
                                   int);
template void Case<double>::turbines_forcing();
template void Case<float>::turbines_forcing();
template void Case<double>::rebuild_turbines();
template void Case<float>::rebuild_turbines();
template void Case<double>::pull_turbine_state();
template void Case<float>::pull_turbine_state();

}  // namespace lg
