// Pointwise / stencil kernels of the time step.  Each kernel cites the
// reference loop it replaces.  Index (k, j, i) of a field lives at
// ((k * ny) + j) * ld + i; k = 0..nz (lbz = 0, the one-rank PPMPI layout).
#pragma once

#include "clock.hpp"
#include "common.hpp"

namespace lg {

// A launch covers levels k0..k0+nlev-1 of every member: grid.y = nlev * M,
// member m = blockIdx.y / nlev; member m's arrays start at m * mstride.
struct Geo {
    int nx, ny, nz, ld;
    long plane;
    int k0, nlev;
    long mstride;
};

__device__ __forceinline__ bool ij_of(int W, int ny, int& j, int& i) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= W * ny) return false;
    j = idx / W;
    i = idx - j * W;
    return true;
}

__device__ __forceinline__ bool pt_of(const Geo& g, int W, int& m, int& k, int& j, int& i,
                                      long& mb) {
    if (!ij_of(W, g.ny, j, i)) return false;
    m = blockIdx.y / g.nlev;
    k = g.k0 + (int)(blockIdx.y - m * g.nlev);
    mb = (long)m * g.mstride;
    return true;
}

// f[m * mstride + q] = v for q < n, every member (grid.y = members)
template <typename T>
__global__ void k_fill_planes(T* f, long mstride, long n, T v) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) f[(long)blockIdx.y * mstride + idx] = v;
}

// Up to 16 (plane, value) fills in one launch: grid (blocks per plane,
// members, entries); entry q sets p[q][m * mstride + i] = v[q], i < plane.
constexpr int MAX_FILLS = 16;
template <typename T>
struct FillList {
    T* p[MAX_FILLS];
    T v[MAX_FILLS];
    long mstride, plane;
};

template <typename T>
__global__ void k_fill_list(FillList<T> a) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.plane) return;
    a.p[blockIdx.z][(long)blockIdx.y * a.mstride + idx] = a.v[blockIdx.z];
}

// f[m * mstride + q] = v for q < n on the members with flag[m] != 0
template <typename T>
__global__ void k_fill_members_if(T* f, const int* flag, long mstride, long n, T v) {
    if (!flag[blockIdx.y]) return;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) f[(long)blockIdx.y * mstride + idx] = v;
}

template <typename T>
__global__ void k_copy_plane(const T* src, T* dst, long mstride, long n) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) dst[(long)blockIdx.y * mstride + idx] = src[(long)blockIdx.y * mstride + idx];
}

// ddz_uv (derivatives.f90:214-264): dfdz(k) = (1/dz)(f(k) - f(k-1)), k=1..nz.
template <typename T>
__global__ void k_ddz_uv(const T* f, T* out, Geo g, T rdz) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.nx, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    out[o] = rdz * (f[o] - f[o - g.plane]);
}

// ddz_w (derivatives.f90:267-311): dfdz(k) = (1/dz)(f(k+1) - f(k)), k=0..nz-1.
template <typename T>
__global__ void k_ddz_w(const T* f, T* out, Geo g, T rdz) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.nx, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    out[o] = rdz * (f[o + g.plane] - f[o]);
}

// Two consecutive values of a row (i even; rows and fields hold an even
// number of values, so the pair is aligned to its size).
template <typename T> struct Pair2;
template <> struct Pair2<float> { using type = float2; };
template <> struct Pair2<double> { using type = double2; };

template <typename T>
__device__ __forceinline__ void ld_pair(const T* p, T& a, T& b) {
    const auto v = *reinterpret_cast<const typename Pair2<T>::type*>(p);
    a = v.x;
    b = v.y;
}

// both values, or the first alone when the second is past the row's end
template <typename T>
__device__ __forceinline__ void st_pair(T* p, T a, T b, bool both) {
    if (both) {
        typename Pair2<T>::type v;
        v.x = a;
        v.y = b;
        *reinterpret_cast<typename Pair2<T>::type*>(p) = v;
    } else {
        p[0] = a;
    }
}

// ddz_uv of u and v and ddz_w of w in one pass (derivatives.f90:214-311),
// the same expressions as k_ddz_uv / k_ddz_w: each thread takes two
// consecutive points (i, i + 1) of a row and DDZ_L levels of them into
// registers, so every plane of u, v, w is read once, in pairs.  Written:
// dudz, dvdz on levels 2..nz-1, dwdz on 1..nz-1; the other levels the two
// kernels compute are overwritten afterwards (BOGUS at 0, 1, nz of dudz /
// dvdz and 0, nz of dwdz) and are set by the caller.
constexpr int DDZ_L = 8;
template <typename T>
__global__ void k_ddz_uvw(const T* u, const T* v, const T* w, T* dudz, T* dvdz, T* dwdz, Geo g,
                          T rdz) {
    int j, ip;
    if (!ij_of((g.nx + 1) / 2, g.ny, j, ip)) return;
    const int i = 2 * ip;
    const bool both = i + 1 < g.nx;
    const int nch = (g.nz - 1 + DDZ_L - 1) / DDZ_L;
    const int m = blockIdx.y / nch;
    const int kb = 1 + (int)(blockIdx.y - m * nch) * DDZ_L;
    const int ke = min(kb + DDZ_L - 1, g.nz - 1);
    const long P = g.plane;
    const long base = (long)m * g.mstride + (long)j * g.ld + i;
    // every level of the chunk is loaded before the first store (level
    // indices clamped into the chunk, so no load depends on a branch)
    T U[2][DDZ_L + 1], V[2][DDZ_L + 1], W[2][DDZ_L + 1];
#pragma unroll
    for (int q = 0; q <= DDZ_L; ++q) {
        const long ouv = base + (long)min(kb - 1 + q, ke) * P;
        const long ow = base + (long)min(kb + q, ke + 1) * P;
        ld_pair(u + ouv, U[0][q], U[1][q]);
        ld_pair(v + ouv, V[0][q], V[1][q]);
        ld_pair(w + ow, W[0][q], W[1][q]);
    }
#pragma unroll
    for (int q = 0; q < DDZ_L; ++q) {
        const int k = kb + q;
        if (k <= ke) {
            const long o = base + (long)k * P;
            if (k >= 2) {
                st_pair(dudz + o, rdz * (U[0][q + 1] - U[0][q]), rdz * (U[1][q + 1] - U[1][q]), both);
                st_pair(dvdz + o, rdz * (V[0][q + 1] - V[0][q]), rdz * (V[1][q + 1] - V[1][q]), both);
            }
            st_pair(dwdz + o, rdz * (W[0][q + 1] - W[0][q]), rdz * (W[1][q + 1] - W[1][q]), both);
        }
    }
}

// ws_equilibrium_lbc (wallstress.f90:171-219) on the k=1 plane; u1f, v1f
// are the test-filtered k=1 planes (member stride fmstride), the other
// pointers field bases; denom[m] = ln(0.5 dz / z0_m).
template <typename T>
__global__ void k_wall_lbc(const T* u1f, const T* v1f, long fmstride, const T* u, const T* v,
                           T* txz, T* tyz, T* dudz, T* dvdz, Geo g, const double* denom,
                           double vonk, double dz) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.nx, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    const long of = (long)m * fmstride + (long)j * g.ld + i;
    const double uf = u1f[of], vf = v1f[of];
    const double ua = sqrt(uf * uf + vf * vf);
    const double us = ua * vonk / denom[m];
    const double c = -(us * us) / ua;
    txz[o] = (T)(c * uf);
    tyz[o] = (T)(c * vf);
    const double uu = u[o], vv = v[o];
    const double a = us / (0.5 * dz * vonk);
    dudz[o] = (uu == 0.0) ? (T)0 : (T)(a * uu / ua);
    dvdz[o] = (vv == 0.0) ? (T)0 : (T)(a * vv / ua);
}

// calc_Sij (sgs_stag_util.f90:467-634), one rank, lbc_mom >= 1, ubc_mom = 0.
template <typename T>
struct SijArgs {
    const T *dudx, *dudy, *dudz, *dvdx, *dvdy, *dvdz, *dwdx, *dwdy, *dwdz;
    T *S11, *S12, *S13, *S22, *S23, *S33;
};

// S_ij at point o of level k (calc_Sij, sgs_stag_util.f90:436-515 with the
// w-node interpolations of the one-rank case)
template <typename T>
struct Sij6 {
    T s11, s12, s13, s22, s23, s33;
};

template <typename T>
__device__ __forceinline__ Sij6<T> sij_at(const SijArgs<T>& a, long o, long P, int k, int nz) {
    // no FMA contraction: every pass computes S_ij, Nu_t and tau_ij with the
    // same roundings, so members batched with others take the same bits
    // whichever pass a step uses
#pragma clang fp contract(off)
    const T h = (T)0.5;
    Sij6<T> r;
    if (k == 1) {
        r.s11 = a.dudx[o];
        r.s12 = h * (a.dudy[o] + a.dvdx[o]);
        const T wx = h * (a.dwdx[o] + a.dwdx[o + P]);
        r.s13 = h * (a.dudz[o] + wx);
        r.s22 = a.dvdy[o];
        const T wy = h * (a.dwdy[o] + a.dwdy[o + P]);
        r.s23 = h * (a.dvdz[o] + wy);
        r.s33 = a.dwdz[o];
    } else if (k == nz) {
        const long m = o - P;
        r.s11 = a.dudx[m];
        r.s12 = h * (a.dudy[m] + a.dvdx[m]);
        r.s13 = h * (a.dudz[o] + a.dwdx[o]);
        r.s22 = a.dvdy[m];
        r.s23 = h * (a.dvdz[o] + a.dwdy[o]);
        r.s33 = h * (a.dwdz[m] + (T)0);
    } else {
        const long m = o - P;
        r.s11 = h * (a.dudx[o] + a.dudx[m]);
        const T uy = a.dudy[o] + a.dudy[m];
        const T vx = a.dvdx[o] + a.dvdx[m];
        r.s12 = (T)0.25 * (uy + vx);
        r.s13 = h * (a.dudz[o] + a.dwdx[o]);
        r.s22 = h * (a.dvdy[o] + a.dvdy[m]);
        r.s23 = h * (a.dvdz[o] + a.dwdy[o]);
        r.s33 = h * (a.dwdz[o] + a.dwdz[m]);
    }
    return r;
}

template <typename T>
__global__ void k_calc_sij(SijArgs<T> a, Geo g) {
    // no FMA contraction: every pass computes S_ij, Nu_t and tau_ij with the
    // same roundings, so members batched with others take the same bits
    // whichever pass a step uses
#pragma clang fp contract(off)
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.nx, m, k, j, i, mb)) return;
    const long P = g.plane;
    const long o = mb + (long)k * P + (long)j * g.ld + i;
    const Sij6<T> r = sij_at(a, o, P, k, g.nz);
    a.S11[o] = r.s11;
    a.S12[o] = r.s12;
    a.S13[o] = r.s13;
    a.S22[o] = r.s22;
    a.S23[o] = r.s23;
    a.S33[o] = r.s33;
}

template <typename T>
__device__ __forceinline__ T smag_norm(T s11, T s12, T s13, T s22, T s23, T s33) {
    // no FMA contraction: every pass computes S_ij, Nu_t and tau_ij with the
    // same roundings, so members batched with others take the same bits
    // whichever pass a step uses
#pragma clang fp contract(off)
    return sqrt((T)2 * (s11 * s11 + s22 * s22 + s33 * s33 +
                        (T)2 * (s12 * s12 + s13 * s13 + s23 * s23)));
}

// sgs_stag: Nu_t and tau_ij (sgs_stag_util.f90:226-235, 237-428), one rank.
template <typename T>
struct TauArgs {
    const T *S11, *S12, *S13, *S22, *S23, *S33, *Cs;
    T *Nu_t, *txx, *txy, *tyy, *txz, *tyz, *tzz;
    T l2, nu;
};

template <typename T>
__device__ __forceinline__ T nut_at(const TauArgs<T>& a, long o) {
    // no FMA contraction: every pass computes S_ij, Nu_t and tau_ij with the
    // same roundings, so members batched with others take the same bits
    // whichever pass a step uses
#pragma clang fp contract(off)
    return smag_norm(a.S11[o], a.S12[o], a.S13[o], a.S22[o], a.S23[o], a.S33[o]) * a.Cs[o] * a.l2;
}

template <typename T>
__global__ void k_sgs_tau(TauArgs<T> a, Geo g) {
    // no FMA contraction: every pass computes S_ij, Nu_t and tau_ij with the
    // same roundings, so members batched with others take the same bits
    // whichever pass a step uses
#pragma clang fp contract(off)
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.nx, m, k, j, i, mb)) return;
    const long P = g.plane;
    const long o = mb + (long)k * P + (long)j * g.ld + i;
    const T nk = nut_at(a, o);
    a.Nu_t[o] = nk;
    const T nu = a.nu;
    if (k == 1) {
        const T c = (T)-2 * (nk + nu);
        a.txx[o] = c * a.S11[o];
        a.txy[o] = c * a.S12[o];
        a.tyy[o] = c * a.S22[o];
        a.tzz[o] = c * a.S33[o];
    } else if (k == g.nz - 1) {
        const long q = o + P;
        const T nk1 = nut_at(a, q);
        const T c = (T)0.5 * (nk + nk1) + nu;
        const T c2 = (T)2 * (nk + nu);
        a.txx[o] = -c * (a.S11[o] + a.S11[q]);
        a.txy[o] = -c * (a.S12[o] + a.S12[q]);
        a.tyy[o] = -c * (a.S22[o] + a.S22[q]);
        a.tzz[o] = -c * (a.S33[o] + a.S33[q]);
        a.txz[o] = -c2 * a.S13[o];
        a.tyz[o] = -c2 * a.S23[o];
    } else if (k < g.nz - 1) {
        const long q = o + P;
        const T nk1 = nut_at(a, q);
        const T c = (T)-0.5 * (nk + nk1);
        const T c2 = (T)-2 * nk;
        const T c3 = (T)-2 * nu * (T)0.5;
        const T c4 = (T)-2 * nu;
        a.txx[o] = (c + c3) * (a.S11[o] + a.S11[q]);
        a.txy[o] = (c + c3) * (a.S12[o] + a.S12[q]);
        a.tyy[o] = (c + c3) * (a.S22[o] + a.S22[q]);
        a.tzz[o] = (c + c3) * (a.S33[o] + a.S33[q]);
        a.txz[o] = (c2 + c4) * a.S13[o];
        a.tyz[o] = (c2 + c4) * a.S23[o];
    }
}

// Nu_t alone (sgs_stag_util.f90:226-235), from stored S_ij: nut_at, as in
// k_sgs_tau.
template <typename T>
__global__ void k_nut(TauArgs<T> a, Geo g) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.nx, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    a.Nu_t[o] = nut_at(a, o);
}

// The nine derivatives at one point of one level.
template <typename T>
struct Der9 {
    T ux, uy, uz, vx, vy, vz, wx, wy, wz;
};

template <typename T>
__device__ __forceinline__ Der9<T> der_at(const SijArgs<T>& a, long o) {
    return Der9<T>{a.dudx[o], a.dudy[o], a.dudz[o], a.dvdx[o], a.dvdy[o],
                   a.dvdz[o], a.dwdx[o], a.dwdy[o], a.dwdz[o]};
}

// sij_at's formulas on derivatives held in registers: dm at k - 1, d at k,
// dp at k + 1 (read for k = 1 only).
template <typename T>
__device__ __forceinline__ Sij6<T> sij_of(const Der9<T>& dm, const Der9<T>& d,
                                          const Der9<T>& dp, int k, int nz) {
    // no FMA contraction: every pass computes S_ij, Nu_t and tau_ij with the
    // same roundings, so members batched with others take the same bits
    // whichever pass a step uses
#pragma clang fp contract(off)
    const T h = (T)0.5;
    Sij6<T> r;
    if (k == 1) {
        r.s11 = d.ux;
        r.s12 = h * (d.uy + d.vx);
        const T wx = h * (d.wx + dp.wx);
        r.s13 = h * (d.uz + wx);
        r.s22 = d.vy;
        const T wy = h * (d.wy + dp.wy);
        r.s23 = h * (d.vz + wy);
        r.s33 = d.wz;
    } else if (k == nz) {
        r.s11 = dm.ux;
        r.s12 = h * (dm.uy + dm.vx);
        r.s13 = h * (d.uz + d.wx);
        r.s22 = dm.vy;
        r.s23 = h * (d.vz + d.wy);
        r.s33 = h * (dm.wz + (T)0);
    } else {
        r.s11 = h * (d.ux + dm.ux);
        const T uy = d.uy + dm.uy;
        const T vx = d.vx + dm.vx;
        r.s12 = (T)0.25 * (uy + vx);
        r.s13 = h * (d.uz + d.wx);
        r.s22 = h * (d.vy + dm.vy);
        r.s23 = h * (d.vz + d.wy);
        r.s33 = h * (d.wz + dm.wz);
    }
    return r;
}

// calc_Sij and the stresses in one pass, for steps without a dynamic-model
// update in between.  A thread walks SIJ_L consecutive levels of one (i, j)
// and keeps the level below's derivatives and S_ij in registers, so each
// level's nine derivatives and Cs_opt2 are read once; it writes S_ij (as
// k_calc_sij does), Nu_t and tau_ij (as k_sgs_tau does, same expressions);
// with store_s = 0 only tau_ij (S_ij and Nu_t are made later by k_calc_sij
// and k_nut, to the same bits).  Grid: x over (i, j), y over chunks of
// SIJ_L levels times members.
constexpr int SIJ_L = 8;

template <typename T>
__global__ void k_sij_tau(SijArgs<T> sa, TauArgs<T> a, Geo g, int store_s) {
    // no FMA contraction: every pass computes S_ij, Nu_t and tau_ij with the
    // same roundings, so members batched with others take the same bits
    // whichever pass a step uses
#pragma clang fp contract(off)
    int j, i;
    if (!ij_of(g.nx, g.ny, j, i)) return;
    const int nch = (g.nz + SIJ_L - 1) / SIJ_L;
    const int m = blockIdx.y / nch;
    const int kb = 1 + (int)(blockIdx.y - m * nch) * SIJ_L;
    const int ke = min(kb + SIJ_L - 1, g.nz);
    const long P = g.plane;
    const long base = (long)m * g.mstride + (long)j * g.ld + i;
    const int nz = g.nz;
    const T nu = a.nu;
    // derivatives at kb - 1 (when kb >= 2), kb and kb + 1
    Der9<T> dm{}, d = der_at(sa, base + (long)kb * P), dp{};
    if (kb >= 2) dm = der_at(sa, base + (long)(kb - 1) * P);
    if (kb < nz) dp = der_at(sa, base + (long)(kb + 1) * P);
    Sij6<T> s0 = sij_of(dm, d, dp, kb, nz);
    T nk = smag_norm(s0.s11, s0.s12, s0.s13, s0.s22, s0.s23, s0.s33) * a.Cs[base + (long)kb * P] * a.l2;
    // level k + 1's derivatives and Cs_opt2 are in registers when level k is
    // worked on; level k + 2's are loaded before level k's stores
    Der9<T> d1 = dp;
    T cs1 = kb < nz ? a.Cs[base + (long)(kb + 1) * P] : (T)0;
#pragma unroll
    for (int q = 0; q < SIJ_L; ++q) {
        const int k = kb + q;
        if (k <= ke) {
            const long o = base + (long)k * P;
            Der9<T> d2{};
            T cs2 = (T)0;
            if (k + 2 <= min(ke + 1, nz)) {
                d2 = der_at(sa, o + 2 * P);
                cs2 = a.Cs[o + 2 * P];
            }
            if (store_s) {
                sa.S11[o] = s0.s11;
                sa.S12[o] = s0.s12;
                sa.S13[o] = s0.s13;
                sa.S22[o] = s0.s22;
                sa.S23[o] = s0.s23;
                sa.S33[o] = s0.s33;
                a.Nu_t[o] = nk;
            }
            if (k == 1) {
                const T c = (T)-2 * (nk + nu);
                a.txx[o] = c * s0.s11;
                a.txy[o] = c * s0.s12;
                a.tyy[o] = c * s0.s22;
                a.tzz[o] = c * s0.s33;
            }
            if (k <= nz - 1) {
                // S_ij and Nu_t at k + 1 from the derivatives at k and k + 1
                // (k + 2 is not read: k + 1 >= 2)
                const Der9<T> dnext = d1;
                const Sij6<T> s1 = sij_of(d, dnext, dnext, k + 1, nz);
                const T nk1 = smag_norm(s1.s11, s1.s12, s1.s13, s1.s22, s1.s23, s1.s33) *
                              cs1 * a.l2;
                if (k == nz - 1 && k != 1) {
                    const T c = (T)0.5 * (nk + nk1) + nu;
                    const T c2 = (T)2 * (nk + nu);
                    a.txx[o] = -c * (s0.s11 + s1.s11);
                    a.txy[o] = -c * (s0.s12 + s1.s12);
                    a.tyy[o] = -c * (s0.s22 + s1.s22);
                    a.tzz[o] = -c * (s0.s33 + s1.s33);
                    a.txz[o] = -c2 * s0.s13;
                    a.tyz[o] = -c2 * s0.s23;
                } else if (k != 1) {
                    const T c = (T)-0.5 * (nk + nk1);
                    const T c2 = (T)-2 * nk;
                    const T c3 = (T)-2 * nu * (T)0.5;
                    const T c4 = (T)-2 * nu;
                    a.txx[o] = (c + c3) * (s0.s11 + s1.s11);
                    a.txy[o] = (c + c3) * (s0.s12 + s1.s12);
                    a.tyy[o] = (c + c3) * (s0.s22 + s1.s22);
                    a.tzz[o] = (c + c3) * (s0.s33 + s1.s33);
                    a.txz[o] = (c2 + c4) * s0.s13;
                    a.tyz[o] = (c2 + c4) * s0.s23;
                }
                d = dnext;
                s0 = s1;
                nk = nk1;
            }
            d1 = d2;
            cs1 = cs2;
        }
    }
}

// divstress_uv sum (divstress_uv.f90:58-83): k = 1..nz-1, all ld columns;
// dxy holds ddx(t_x1) + ddy(t_x2) formed in spectral space.
template <typename T>
__global__ void k_divt_uv(const T* dxy, const T* tz, T* out, Geo g, T rdz) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.ld, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    if (i >= g.nx) {
        out[o] = (T)0;
        return;
    }
    const T dz = rdz * (tz[o + g.plane] - tz[o]);
    out[o] = dxy[o] + dz;
}

// divstress_w sum (divstress_w.f90:63-114): k = 1..nz.
template <typename T>
__global__ void k_divt_w(const T* dxy, const T* tzz, T* out, Geo g, T rdz) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.ld, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    if (i >= g.nx) {
        if (k < g.nz) out[o] = (T)0;
        return;
    }
    if (k == 1 || k == g.nz) {
        out[o] = dxy[o];
    } else {
        const T dz = rdz * (tzz[o] - tzz[o - g.plane]);
        out[o] = dxy[o] + dz;
    }
}

// convec vorticity (convec.f90:97-168), scaled by 1/(nx ny), k = 1..nz.
template <typename T>
struct VortArgs {
    const T *dudy, *dudz, *dvdx, *dvdz, *dwdx, *dwdy;
    T *o1, *o2, *o3;
    T c;
};

template <typename T>
__global__ void k_vort(VortArgs<T> a, Geo g) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.nx, m, k, j, i, mb)) return;
    const long P = g.plane;
    const long o = mb + (long)k * P + (long)j * g.ld + i;
    if (k == 1) {
        a.o1[o] = a.c * ((T)0.5 * (a.dwdy[o] + a.dwdy[o + P]) - a.dvdz[o]);
        a.o2[o] = a.c * (a.dudz[o] - (T)0.5 * (a.dwdx[o] + a.dwdx[o + P]));
    } else {
        a.o1[o] = a.c * (a.dwdy[o] - a.dvdz[o]);
        a.o2[o] = a.c * (a.dudz[o] - a.dwdx[o]);
    }
    a.o3[o] = a.c * (a.dvdx[o] - a.dudy[o]);
}

// Vorticity in spectral space from the filtered u, v, w spectra (normalised
// by 1/(nx ny) like the reference's const * vort before its r2c,
// convec.f90:97-168): w1 = i ky w - dz v, w2 = dz u - i kx w, w3 = i kx v -
// i ky u; the wall level uses the averaged dw/dx, dw/dy and the wall-model
// dudz, dvdz planes (Du1, Dv1), the lid level the zero dudz, dvdz of
// ws_free_ubc.  k = 1..nz.
template <typename T>
__global__ void k_vort_spec(const cplx<T>* U, const cplx<T>* V, const cplx<T>* W,
                            const cplx<T>* Dv1, const cplx<T>* Du1, cplx<T>* O1, cplx<T>* O2,
                            cplx<T>* O3, int ny, int NK, int nz, long splane, T kxs, T kys,
                            T rdz, long smstride) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ny * NK) return;
    const int m = blockIdx.y / nz;
    const int k = blockIdx.y - m * nz + 1;
    const long mb = (long)m * smstride;
    U += mb;
    V += mb;
    W += mb;
    Dv1 += mb;
    Du1 += mb;
    O1 += mb;
    O2 += mb;
    O3 += mb;
    const int j = idx / NK, c = idx - j * NK;
    T kx = kxs * (T)c, ky = kys * (T)((j < ny / 2) ? j : j - ny);
    if (j == ny / 2) kx = ky = (T)0;
    auto muli = [](cplx<T> a, T kk) { return cplx<T>{-a.y * kk, a.x * kk}; };
    const long o = (long)k * splane + idx;
    const cplx<T> w = W[o];
    cplx<T> o1, o2;
    if (k == 1) {
        const cplx<T> w2 = W[o + splane];
        const cplx<T> ay = muli(w, ky), by = muli(w2, ky);
        const cplx<T> ax = muli(w, kx), bx = muli(w2, kx);
        o1 = cplx<T>{(T)0.5 * (ay.x + by.x) - Dv1[idx].x, (T)0.5 * (ay.y + by.y) - Dv1[idx].y};
        o2 = cplx<T>{Du1[idx].x - (T)0.5 * (ax.x + bx.x), Du1[idx].y - (T)0.5 * (ax.y + bx.y)};
    } else if (k == nz) {
        o1 = muli(w, ky);
        const cplx<T> ax = muli(w, kx);
        o2 = cplx<T>{-ax.x, -ax.y};
    } else {
        const cplx<T> v0 = V[o], vm = V[o - splane];
        const cplx<T> u0 = U[o], um = U[o - splane];
        const cplx<T> ay = muli(w, ky), ax = muli(w, kx);
        o1 = cplx<T>{ay.x - rdz * (v0.x - vm.x), ay.y - rdz * (v0.y - vm.y)};
        o2 = cplx<T>{rdz * (u0.x - um.x) - ax.x, rdz * (u0.y - um.y) - ax.y};
    }
    const cplx<T> vx = muli(V[o], kx), uy = muli(U[o], ky);
    O1[o] = o1;
    O2[o] = o2;
    O3[o] = cplx<T>{vx.x - uy.x, vx.y - uy.y};
}

// convec products on the 3/2 grid (convec.f90:172-305), k = 1..nz.
template <typename T>
struct ProdArgs {
    const T *ub, *vb, *wb, *o1, *o2, *o3;
    T *cx, *cy, *cz;
    double c;
};

// Levels 1..nz; each thread takes two consecutive points (i, i + 1) of one
// row and PROD_L levels of them (and the levels either side) into
// registers, so every plane is read once, in pairs.  The products are
// formed in FP64 from the stored values and rounded once to T: their sums
// cancel where the flow has sharp peaks, and float arithmetic there would
// add errors the velocity field itself does not carry.  No FMA contraction:
// the rounding does not depend on the compiler's fusion choices.
constexpr int PROD_L = 8;
template <typename T>
__global__ void __launch_bounds__(256) k_products(ProdArgs<T> a, int nx2, int ny2, int ldb, int nz, long bmstride) {
#pragma clang fp contract(off)
    int j, ip;
    const int nxp = (nx2 + 1) / 2;
    if (!ij_of(nxp, ny2, j, ip)) return;
    const int i = 2 * ip;
    const bool both = i + 1 < nx2;
    const int nch = (nz + PROD_L - 1) / PROD_L;
    const int m = blockIdx.y / nch;
    const int kb = 1 + (int)(blockIdx.y - m * nch) * PROD_L;
    const int ke = min(kb + PROD_L - 1, nz);
    const long P = (long)ny2 * ldb;
    const long base = (long)m * bmstride + (long)j * ldb + i;
    const double h = 0.5;
    const double c = a.c;
    // every level of the chunk is loaded before the first store (level
    // indices clamped into range, so no load depends on a branch): ub, vb,
    // o3 at kb + q (q < PROD_L) and kb - 1; wb, o1, o2 at kb + q (q <= PROD_L)
    T U[2][PROD_L], V[2][PROD_L], O3[2][PROD_L], W[2][PROD_L + 1], O1[2][PROD_L + 1],
        O2[2][PROD_L + 1];
    T ubm0[2], vbm0[2];
    {
        const long om = base + (long)max(kb - 1, 0) * P;
        ld_pair(a.ub + om, ubm0[0], ubm0[1]);
        ld_pair(a.vb + om, vbm0[0], vbm0[1]);
    }
#pragma unroll
    for (int q = 0; q <= PROD_L; ++q) {
        const long ow = base + (long)min(kb + q, nz) * P;
        ld_pair(a.wb + ow, W[0][q], W[1][q]);
        ld_pair(a.o1 + ow, O1[0][q], O1[1][q]);
        ld_pair(a.o2 + ow, O2[0][q], O2[1][q]);
        if (q < PROD_L) {
            const long ou = base + (long)min(kb + q, ke) * P;
            ld_pair(a.ub + ou, U[0][q], U[1][q]);
            ld_pair(a.vb + ou, V[0][q], V[1][q]);
            ld_pair(a.o3 + ou, O3[0][q], O3[1][q]);
        }
    }
    // per point: ub, vb at k - 1; wb, o1, o2 at k
    double ubm[2], vbm[2], wb0[2], o10[2], o20[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        ubm[e] = kb >= 2 ? (double)ubm0[e] : 0.0;
        vbm[e] = kb >= 2 ? (double)vbm0[e] : 0.0;
        wb0[e] = W[e][0];
        o10[e] = O1[e][0];
        o20[e] = O2[e][0];
    }
#pragma unroll
    for (int q = 0; q < PROD_L; ++q) {
        const int k = kb + q;
        if (k <= ke) {
            const long o = base + (long)k * P;
            double x[2] = {0.0, 0.0}, y[2] = {0.0, 0.0}, z[2] = {0.0, 0.0};
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const double ub0 = U[e][q], vb0 = V[e][q];
                double wb1 = 0.0, o11 = 0.0, o21 = 0.0;
                if (k <= nz - 1) {
                    wb1 = W[e][q + 1];
                    o11 = O1[e][q + 1];
                    o21 = O2[e][q + 1];
                    const double o30 = O3[e][q];
                    if (k == 1) {
                        x[e] = c * (vb0 * (-o30) + h * wb1 * o21);
                        y[e] = c * (ub0 * o30 + h * wb1 * (-o11));
                    } else if (k == nz - 1) {
                        x[e] = c * (vb0 * (-o30) + h * wb0[e] * o20[e]);
                        y[e] = c * (ub0 * o30 + h * wb0[e] * (-o10[e]));
                    } else {
                        x[e] = c * (vb0 * (-o30) + h * (wb1 * o21 + wb0[e] * o20[e]));
                        y[e] = c * (ub0 * o30 + h * (wb1 * (-o11) + wb0[e] * (-o10[e])));
                    }
                }
                if (k >= 2 && k <= nz - 1)
                    z[e] = (c * h) * ((ub0 + ubm[e]) * (-o20[e]) + (vb0 + vbm[e]) * o10[e]);
                ubm[e] = ub0;
                vbm[e] = vb0;
                wb0[e] = wb1;
                o10[e] = o11;
                o20[e] = o21;
            }
            if (k <= nz - 1) {
                st_pair(a.cx + o, (T)x[0], (T)x[1], both);
                st_pair(a.cy + o, (T)y[0], (T)y[1], both);
            }
            st_pair(a.cz + o, (T)z[0], (T)z[1], both);
        }
    }
}

// RHS assembly, forcing and AB2 (main.f90:211-214, 228-232, 264-266, 287-296);
// members with euler[m] take the first-step Euler update with RHS_f = RHS
// (main.f90:273-280).
template <typename T>
struct Ab2Args {
    T *u, *v, *w, *rx, *ry, *rz;
    T *rxf, *ryf, *rzf;
    const T *dx, *dy, *dz, *fx, *fy, *fz;
    T dt, t1, t2;
    const double* mpf;     // 2 per member: mean_p_force_x, mean_p_force_y
    const int* euler;      // per member
    int use_mpf;
    const double* clk;     // per-member clock (CFL-adaptive dt) or null: dt, t1, t2
};

template <typename T>
__global__ void k_rhs_ab2(Ab2Args<T> a, Geo g) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.ld, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    T dt = a.dt, t1 = a.t1, t2 = a.t2;
    if (a.clk) {
        const double* c = a.clk + (long)CK_N * m;
        dt = (T)c[CK_DT];
        t1 = (T)c[CK_TADV1];
        t2 = (T)c[CK_TADV2];
    }
    if (k <= g.nz - 1) {
        T rx = -a.rx[o] - a.dx[o];
        T ry = -a.ry[o] - a.dy[o];
        T rz = -a.rz[o] - a.dz[o];
        if (a.use_mpf) {
            rx = rx + (T)a.mpf[2 * m];
            ry = ry + (T)a.mpf[2 * m + 1];
        }
        if (a.fx) {
            rx = rx + a.fx[o];
            ry = ry + a.fy[o];
            rz = rz + a.fz[o];
        }
        a.rx[o] = rx;
        a.ry[o] = ry;
        a.rz[o] = rz;
        T fx = rx, fy = ry, fz = rz;
        if (a.euler[m]) {
            a.rxf[o] = rx;
            a.ryf[o] = ry;
            a.rzf[o] = rz;
        } else {
            fx = a.rxf[o];
            fy = a.ryf[o];
            fz = a.rzf[o];
        }
        a.u[o] = a.u[o] + dt * (t1 * rx + t2 * fx);
        a.v[o] = a.v[o] + dt * (t1 * ry + t2 * fy);
        a.w[o] = a.w[o] + dt * (t1 * rz + t2 * fz);
    } else {
        const T rz = -a.rz[o] - a.dz[o];
        a.rz[o] = rz;
        T fz = rz;
        if (a.euler[m]) {
            a.rzf[o] = rz;
            a.rxf[o] = a.rx[o];
            a.ryf[o] = a.ry[o];
        } else {
            fz = a.rzf[o];
        }
        a.w[o] = a.w[o] + dt * (t1 * rz + t2 * fz);
    }
}

// RHS -= grad p (main.f90:321-326), k = 1..nz (z only at nz).
template <typename T>
__global__ void k_rhs_gradp(T* rx, T* ry, T* rz, const T* px, const T* py, const T* pz, Geo g) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.ld, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    if (k <= g.nz - 1) {
        rx[o] = rx[o] - px[o];
        ry[o] = ry[o] - py[o];
    }
    rz[o] = rz[o] - pz[o];
}

// dpdz by differences (press_stag_array.f90:284-288), k = 1..nz.
template <typename T>
__global__ void k_dpdz(const T* p, T* dpdz, Geo g, T dzv) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.nx, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    dpdz[o] = (p[o] - p[o - g.plane]) / dzv;
}

// project (forcing.f90:149-244), interior update; lid/wall BCs follow.
template <typename T>
__global__ void k_project(T* u, T* v, T* w, const T* px, const T* py, const T* pz, Geo g, T dt,
                          T t1, const double* clk) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.nx, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    if (clk) {
        dt = (T)clk[(long)CK_N * m + CK_DT];
        t1 = (T)clk[(long)CK_N * m + CK_TADV1];
    }
    u[o] = u[o] + dt * (-(t1 * px[o]));
    v[o] = v[o] + dt * (-(t1 * py[o]));
    if (k >= 2) w[o] = w[o] + dt * (-(t1 * pz[o]));
}

// Block-partial sums for the diagnostics (rmsdiv.f90:39-50, io.f90:115-124),
// reduced by wave shuffles (block_reduce).
template <typename T>
__global__ void k_rmsdiv_part(const T* dudx, const T* dvdy, const T* dwdz, Geo g, double* part) {
    double acc = 0.0;
    const long n = (long)(g.nz - 1) * g.ny * g.nx;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const long k = idx / ((long)g.ny * g.nx) + 1;
        const long r = idx % ((long)g.ny * g.nx);
        const long j = r / g.nx, i = r % g.nx;
        const long o = k * g.plane + j * g.ld + i;
        acc += fabs((double)dudx[o] + (double)dvdy[o] + (double)dwdz[o]);
    }
    acc = block_reduce(acc, RedSum{}, 0.0);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

template <typename T>
__global__ void k_energy_part(const T* u, const T* v, const T* w, Geo g, double* part) {
    double acc = 0.0;
    const long n = (long)(g.nz - 1) * g.ny * g.nx;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const long k = idx / ((long)g.ny * g.nx) + 1;
        const long r = idx % ((long)g.ny * g.nx);
        const long j = r / g.nx, i = r % g.nx;
        const long o = k * g.plane + j * g.ld + i;
        const double tw = 0.5 * ((double)w[o] + (double)w[o + g.plane]);
        acc += (double)u[o] * u[o] + (double)v[o] * v[o] + tw * tw;
    }
    acc = block_reduce(acc, RedSum{}, 0.0);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

template <typename T>
__global__ void k_maxabs_part(const T* f, Geo g, double* part) {
    double acc = 0.0;
    const long n = (long)(g.nz - 1) * g.ny * g.nx;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const long k = idx / ((long)g.ny * g.nx) + 1;
        const long r = idx % ((long)g.ny * g.nx);
        const long j = r / g.nx, i = r % g.nx;
        acc = fmax(acc, fabs((double)f[k * g.plane + j * g.ld + i]));
    }
    acc = block_reduce(acc, RedMax{}, 0.0);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

}  // namespace lg
