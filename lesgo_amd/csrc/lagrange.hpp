// Lagrangian scale-dependent dynamic model kernels (lagrange_Sdep.f90,
// interpolag_Sdep.f90, functions.f90:188-265,349-474).
#pragma once

#include "common.hpp"
#include "pointwise.hpp"

namespace lg {

struct InterpGeo {
    int nx, ny, nz, ld;
    long plane;
    double dx, dy, dz, lx, ly, lz, ldt;
    long mstride;          // elements between members (all arrays)
    const double* clk;     // per-member clock (CFL-adaptive dt: lagran_dt) or null
};

// cell_indx_w for 'i'/'j' (functions.f90:188-237): px is wrapped with
// Fortran MODULO (the wrapped value is also the one used for xdiff).
__device__ __forceinline__ int cell_xy(double& px, double L, double d, int n) {
    double r = fmod(px, L);
    if (r != 0.0 && r < 0.0) r += L;
    px = r;
    const double thresh = 1.0e-9;
    if (fabs(r) / L < thresh) return 1;
    if (fabs(r - L) / L < thresh) return n;
    if (!(r >= 0.0 && r < L)) return 1;     // non-finite input: keep the gather in bounds
    return min(max((int)floor(r / d) + 1, 1), n);
}

// trilinear_interp_w (functions.f90:349-474) for lbc_mom > 0, ubc_mom = 0,
// split into the stencil of a point (cell corners and offsets, found once)
// and its application to each field.
struct TriStencil {
    long o00, o01, o10, o11;   // (j0, i0), (j0, i1), (j1, i0), (j1, i1) at level k0
    long dk;                   // plane offset from k0 to k1
    double xdiff, ydiff, zdiff;
};

__device__ __forceinline__ TriStencil tri_stencil(const InterpGeo& g, double px, double py,
                                                  double pz) {
    const int i0 = cell_xy(px, g.lx, g.dx, g.nx);
    const int j0 = cell_xy(py, g.ly, g.dy, g.ny);
    const int i1 = (i0 + 1 > g.nx) ? 1 : i0 + 1;
    const int j1 = (j0 + 1 > g.ny) ? 1 : j0 + 1;
    TriStencil t;
    t.xdiff = px - (double)(i0 - 1) * g.dx;
    t.ydiff = py - (double)(j0 - 1) * g.dy;
    const double z1 = (1 - 0.5) * g.dz;
    const double zw2 = ((2 - 0.5) * g.dz) - g.dz / 2.0;
    int k0, k1;
    if (pz < zw2) {
        if (pz < z1) {
            k0 = 1;
            k1 = 1;
            t.zdiff = 0.0;
        } else {
            k0 = 1;
            k1 = 2;
            t.zdiff = 2.0 * (pz - z1);
        }
    } else {
        const double zw1 = ((1 - 0.5) * g.dz) - g.dz / 2.0;
        const double zwn = ((g.nz - 0.5) * g.dz) - g.dz / 2.0;
        if (fabs(pz - zwn) / g.lz < 1.0e-9) k0 = g.nz - 1;
        else if (!(pz == pz)) k0 = 1;
        else k0 = (int)fmin(fmax(floor((pz - zw1) / g.dz) + 1.0, 0.0), (double)g.nz);
        k1 = k0 + 1;
        t.zdiff = pz - (((k0 - 0.5) * g.dz) - g.dz / 2.0);
    }
    // keep the gather inside the array (the departure point is assumed to
    // stay within one cell, interpolag_Sdep.f90:31-34)
    k0 = min(max(k0, 0), g.nz);
    k1 = min(max(k1, 0), g.nz);
    const long base = (long)k0 * g.plane;
    t.o00 = base + (long)(j0 - 1) * g.ld + (i0 - 1);
    t.o01 = base + (long)(j0 - 1) * g.ld + (i1 - 1);
    t.o10 = base + (long)(j1 - 1) * g.ld + (i0 - 1);
    t.o11 = base + (long)(j1 - 1) * g.ld + (i1 - 1);
    t.dk = (long)(k1 - k0) * g.plane;
    return t;
}

template <typename T>
__device__ __forceinline__ double tri_apply(const T* var, const TriStencil& t,
                                            const InterpGeo& g) {
    const double a00 = (double)var[t.o00], a01 = (double)var[t.o01];
    const double a10 = (double)var[t.o10], a11 = (double)var[t.o11];
    const double b00 = (double)var[t.o00 + t.dk], b01 = (double)var[t.o01 + t.dk];
    const double b10 = (double)var[t.o10 + t.dk], b11 = (double)var[t.o11 + t.dk];
    const double u1 = a00 + t.xdiff * (a01 - a00) / g.dx;
    const double u2 = a10 + t.xdiff * (a11 - a10) / g.dx;
    const double u3 = b00 + t.xdiff * (b01 - b00) / g.dx;
    const double u4 = b10 + t.xdiff * (b11 - b10) / g.dx;
    const double u5 = u1 + t.ydiff * (u2 - u1) / g.dy;
    const double u6 = u3 + t.ydiff * (u4 - u3) / g.dy;
    return u5 + t.zdiff * (u6 - u5) / g.dz;
}

// interpolag_Sdep (interpolag_Sdep.f90:73-241): departure points and the
// four gathers; tF* are copies of F_* taken before the call.
template <typename T>
__global__ void k_interpolag(const T* u, const T* v, const T* w, const T* tLM, const T* tMM,
                             const T* tQN, const T* tNN, T* LM, T* MM, T* QN, T* NN,
                             InterpGeo g, const int* mode) {
    int j, i;
    if (!ij_of(g.nx, g.ny, j, i)) return;
    const int m = blockIdx.y / g.nz;
    if (mode[m] == 0) return;      // member not yet dynamic
    const int k = blockIdx.y - m * g.nz + 1;
    const long mb = (long)m * g.mstride;
    if (g.clk) g.ldt = g.clk[(long)CK_N * m + CK_LAGDT];
    u += mb;
    v += mb;
    w += mb;
    tLM += mb;
    tMM += mb;
    tQN += mb;
    tNN += mb;
    LM += mb;
    MM += mb;
    QN += mb;
    NN += mb;
    const long P = g.plane;
    const long o = (long)k * P + (long)j * g.ld + i;
    const double x = (double)i * g.dx;
    const double y = (double)j * g.dy;
    double px, py, pz;
    if (k == 1) {
        px = x - (double)u[o] * g.ldt;
        py = y - (double)v[o] * g.ldt;
        pz = ((1 - 0.5) * g.dz) - 0.25 * (double)w[o + P] * g.ldt;
    } else if (k == g.nz) {
        px = x - (double)u[o - P] * g.ldt;
        py = y - (double)v[o - P] * g.ldt;
        pz = ((g.nz - 0.5) * g.dz) - g.dz / 2.0;
    } else {
        px = x - 0.5 * ((double)u[o - P] + (double)u[o]) * g.ldt;
        py = y - 0.5 * ((double)v[o - P] + (double)v[o]) * g.ldt;
        pz = (((k - 0.5) * g.dz) - g.dz / 2.0) - (double)w[o] * g.ldt;
    }
    const TriStencil t = tri_stencil(g, px, py, pz);
    LM[o] = (T)tri_apply(tLM, t, g);
    MM[o] = (T)tri_apply(tMM, t, g);
    QN[o] = (T)tri_apply(tQN, t, g);
    NN[o] = (T)tri_apply(tNN, t, g);
}

// Quantities to be test filtered for levels k = k0..k0+nc-1
// (lagrange_Sdep.f90:93-151, 211-240): ubar, vbar, wbar, the six products
// and |S| S_ij.  Output q-major planes of (ny, ld).
template <typename T>
__global__ void k_lag_prep(const T* u, const T* v, const T* w, const T* S11, const T* S12,
                           const T* S13, const T* S22, const T* S23, const T* S33, T* out,
                           Geo g, int k0, int nc, int M) {
    int j, i;
    if (!ij_of(g.nx, g.ny, j, i)) return;
    const int m = blockIdx.y / nc;
    const int c = blockIdx.y - m * nc;
    const int k = k0 + c;
    const long P = g.plane;
    const long o = (long)m * g.mstride + (long)k * P + (long)j * g.ld + i;
    T ub, vb, wb;
    if (k == 1) {
        ub = u[o];
        vb = v[o];
        wb = (T)0.25 * w[o + P];
    } else if (k == g.nz) {
        ub = u[o - P];
        vb = v[o - P];
        wb = (T)0;
    } else {
        ub = (T)0.5 * (u[o] + u[o - P]);
        vb = (T)0.5 * (v[o] + v[o - P]);
        wb = w[o];
    }
    const long Q = (long)M * nc * P;
    const long oc = ((long)m * nc + c) * P + (long)j * g.ld + i;
    out[oc + 0 * Q] = ub;
    out[oc + 1 * Q] = vb;
    out[oc + 2 * Q] = wb;
    out[oc + 3 * Q] = ub * ub;
    out[oc + 4 * Q] = ub * vb;
    out[oc + 5 * Q] = ub * wb;
    out[oc + 6 * Q] = vb * vb;
    out[oc + 7 * Q] = vb * wb;
    out[oc + 8 * Q] = wb * wb;
    const T s11 = S11[o], s12 = S12[o], s13 = S13[o], s22 = S22[o], s23 = S23[o], s33 = S33[o];
    const T S = smag_norm(s11, s12, s13, s22, s23, s33);
    out[oc + 9 * Q] = S * s11;
    out[oc + 10 * Q] = S * s12;
    out[oc + 11 * Q] = S * s13;
    out[oc + 12 * Q] = S * s22;
    out[oc + 13 * Q] = S * s23;
    out[oc + 14 * Q] = S * s33;
}

// Filtered-quantity slots: bar (2 delta) in [0,21), hat (4 delta) in [21,42):
// 0-2 ubar,vbar,wbar; 3-8 products 11,12,13,22,23,33; 9-14 S_ij; 15-20 |S|S_ij.
struct LagUpd {
    double opftdelta, powcoeff, cnst, tf1_2, tf2_2, ldt, zero, expo, betamin;
    const double* clk;     // per-member clock (CFL-adaptive dt: lagran_dt) or null
};

// mode[m]: 0 = member not yet dynamic (untouched), 1 = update, 2 = first
// dynamic call of a fresh start: F_MM = MM, F_LM = 0.03 MM, F_NN = NN,
// F_QN = 0.03 NN before the update (lagrange_Sdep.f90:266-277, 326-337; the
// padding columns of F_* are not kept, nothing reads them).
template <typename T>
__global__ void k_lag_update(const T* filt, T* F_LM, T* F_MM, T* F_QN, T* F_NN, T* Cs,
                             T* Beta, T* Tn_all, Geo g, int k0, int nc, int M, LagUpd p,
                             const int* mode, const double* contr, const T* Fin, long Fq) {
    int j, i;
    if (!ij_of(g.ld, g.ny, j, i)) return;
    const int m = blockIdx.y / nc;
    const int c = blockIdx.y - m * nc;
    const int md = mode[m];
    if (md == 0) return;
    if (p.clk) p.ldt = p.clk[(long)CK_N * m + CK_LAGDT];
    const int k = k0 + c;
    const long P = g.plane;
    const long o = (long)m * g.mstride + (long)k * P + (long)j * g.ld + i;
    if (i >= g.nx) {
        // Cs_opt2(ld-1:ld) = zero (lagrange_Sdep.f90:397-399)
        Cs[o] = (T)p.zero;
        return;
    }
    const long Q = (long)M * nc * P;
    const long oc = ((long)m * nc + c) * P + (long)j * g.ld + i;
    // L_ij M_ij, M_ij M_ij, Q_ij N_ij, N_ij N_ij: from the row pass that
    // formed them (contr: four FP64 fields Q apart) or from the 42 filtered
    // fields
    double LM, MM, QN, NN;
    if (contr) {
        LM = contr[oc];
        MM = contr[oc + Q];
        QN = contr[oc + 2 * Q];
        NN = contr[oc + 3 * Q];
    } else {
        auto B = [&](int q) -> double { return (double)filt[oc + (long)q * Q]; };
        auto H = [&](int q) -> double { return (double)filt[oc + (long)(21 + q) * Q]; };
        // L_ij and Q_ij (lagrange_Sdep.f90:137-167)
        const double ub = B(0), vb = B(1), wb = B(2);
        const double uh = H(0), vh = H(1), wh = H(2);
        const double L11 = B(3) - ub * ub, L12 = B(4) - ub * vb, L13 = B(5) - ub * wb;
        const double L22 = B(6) - vb * vb, L23 = B(7) - vb * wb, L33 = B(8) - wb * wb;
        const double Q11 = H(3) - uh * uh, Q12 = H(4) - uh * vh, Q13 = H(5) - uh * wh;
        const double Q22 = H(6) - vh * vh, Q23 = H(7) - vh * wh, Q33 = H(8) - wh * wh;
        // |S_bar|, |S_hat| (lagrange_Sdep.f90:203-209)
        const double S_bar = smag_norm(B(9), B(10), B(11), B(12), B(13), B(14));
        const double S_hat = smag_norm(H(9), H(10), H(11), H(12), H(13), H(14));
        const double cn = p.cnst;
        // M_ij, N_ij (lagrange_Sdep.f90:241-253)
        const double M11 = cn * (B(15) - p.tf1_2 * S_bar * B(9));
        const double M12 = cn * (B(16) - p.tf1_2 * S_bar * B(10));
        const double M13 = cn * (B(17) - p.tf1_2 * S_bar * B(11));
        const double M22 = cn * (B(18) - p.tf1_2 * S_bar * B(12));
        const double M23 = cn * (B(19) - p.tf1_2 * S_bar * B(13));
        const double M33 = cn * (B(20) - p.tf1_2 * S_bar * B(14));
        const double N11 = cn * (H(15) - p.tf2_2 * S_hat * H(9));
        const double N12 = cn * (H(16) - p.tf2_2 * S_hat * H(10));
        const double N13 = cn * (H(17) - p.tf2_2 * S_hat * H(11));
        const double N22 = cn * (H(18) - p.tf2_2 * S_hat * H(12));
        const double N23 = cn * (H(19) - p.tf2_2 * S_hat * H(13));
        const double N33 = cn * (H(20) - p.tf2_2 * S_hat * H(14));
        // contractions (lagrange_Sdep.f90:256-259)
        LM = L11 * M11 + L22 * M22 + L33 * M33 + 2.0 * (L12 * M12 + L13 * M13 + L23 * M23);
        MM = M11 * M11 + M22 * M22 + M33 * M33 + 2.0 * (M12 * M12 + M13 * M13 + M23 * M23);
        QN = Q11 * N11 + Q22 * N22 + Q33 * N33 + 2.0 * (Q12 * N12 + Q13 * N13 + Q23 * N23);
        NN = N11 * N11 + N22 * N22 + N33 * N33 + 2.0 * (N12 * N12 + N13 * N13 + N23 * N23);
    }
    const double zero = p.zero;
    // 2-delta running averages (lagrange_Sdep.f90:297-315)
    // the interpolated F_* (interpolag): in place, or Fin + q Fq when the
    // interpolation wrote them there
    double fLM, fMM, fQN, fNN;
    if (Fin) {
        fLM = Fin[o];
        fMM = Fin[o + Fq];
        fQN = Fin[o + 2 * Fq];
        fNN = Fin[o + 3 * Fq];
    } else {
        fLM = F_LM[o];
        fMM = F_MM[o];
        fQN = F_QN[o];
        fNN = F_NN[o];
    }
    if (md == 2) {
        fMM = MM;
        fLM = 0.03 * MM;
        fNN = NN;
        fQN = 0.03 * NN;
    }
    // Tn^(-1/8) (the model's exponent) as 1 / sqrt(sqrt(sqrt(Tn))): three
    // square roots and a division, within a few ulp of pow at a fraction
    // of its cost; the same values at 0 (inf) and inf (0)
    auto tpow = [&](double x) {
        return p.powcoeff == -0.125 ? 1.0 / sqrt(sqrt(sqrt(x))) : pow(x, p.powcoeff);
    };
    double Tn = fmax(fLM * fMM, zero);
    Tn = p.opftdelta * tpow(Tn);
    Tn = fmax(zero, Tn);
    double dumfac = p.ldt / Tn;
    double epsi = dumfac / (1.0 + dumfac);
    fLM = epsi * LM + (1.0 - epsi) * fLM;
    fMM = epsi * MM + (1.0 - epsi) * fMM;
    fLM = fmax(zero, fLM);
    const double cs2d = fmax(zero, fLM / (fMM + zero));
    // 4-delta running averages (lagrange_Sdep.f90:333-352)
    Tn = fmax(fQN * fNN, zero);
    Tn = p.opftdelta * tpow(Tn);
    Tn = fmax(zero, Tn);
    dumfac = p.ldt / Tn;
    epsi = dumfac / (1.0 + dumfac);
    fQN = epsi * QN + (1.0 - epsi) * fQN;
    fNN = epsi * NN + (1.0 - epsi) * fNN;
    fQN = fmax(zero, fQN);
    const double cs4d = fmax(zero, fQN / (fNN + zero));
    // beta and Cs_opt2 (lagrange_Sdep.f90:377-404)
    const double ratio = cs4d / cs2d;
    double beta = (p.expo == 1.0) ? ratio : pow(ratio, p.expo);
    if (k == g.nz) beta = 1.0;
    const double betaclip = fmax(beta, p.betamin);
    const double cs = fmax(zero, cs2d / betaclip);
    F_LM[o] = (T)fLM;
    F_MM[o] = (T)fMM;
    F_QN[o] = (T)fQN;
    F_NN[o] = (T)fNN;
    Cs[o] = (T)cs;
    Beta[o] = (T)beta;
    Tn_all[o] = (T)Tn;
}

}  // namespace lg
