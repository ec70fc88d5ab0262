// Spectral-resident step: the pointwise parts of the step after the stress
// and advection transforms, done on the (kx, ky) planes the transforms
// already produced instead of on physical fields.
//
// Every operation between the forward transforms of the stresses and the
// next step's filtered velocity is linear and acts plane by plane: the
// divergence of the stresses (divstress_uv.f90:21-85, divstress_w.f90:21-116),
// the RHS assembly and AB2 update (main.f90:211-296), the pressure gradient
// (press_stag_array.f90:251-288, main.f90:321-326) and the projection with
// its boundary rows (forcing.f90:149-244).  Done in spectral space they need
// no inverse transform of divt, RHS, p, dpdx, dpdy and no forward transform
// of u*, and the projected spectra are the next step's filtered velocity
// (filt_da's forward transform of u is the identity on them: the filter only
// removes the Nyquist modes, which the pressure solve has already zeroed).
// Spectra are in filt_da's normalisation, F(f) / (nx ny); planes hold
// ny x NK modes, kx = kxs c (c < NK = nx/2), ky wrapped, the ky Nyquist row
// carrying k = 0 (fft.f90:130-160).
#pragma once

#include "common.hpp"
#include "pointwise.hpp"

namespace lg {

template <typename T>
struct SpecGeo {
    int ny, NK, nz;
    int k0, nlev;        // levels k0 .. k0 + nlev - 1 (grid.y = nlev * members)
    long splane;         // ny * NK
    long msp;            // complex elements between members
    T kxs, kys;          // 2 pi / Lx, 2 pi / Ly
};

template <typename T>
__device__ __forceinline__ bool spec_of(const SpecGeo<T>& g, int& m, int& k, int& j, int& c,
                                        long& o) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.ny * g.NK) return false;
    m = blockIdx.y / g.nlev;
    k = g.k0 + (blockIdx.y - m * g.nlev);
    j = idx / g.NK;
    c = idx - j * g.NK;
    o = (long)m * g.msp + (long)k * g.splane + idx;
    return true;
}

// i kx, i ky of the mode (j, c) as apply_op forms them (Nyquist row: 0)
template <typename T>
__device__ __forceinline__ void spec_k(const SpecGeo<T>& g, int j, int c, T& kx, T& ky) {
    if (j == g.ny / 2) {
        kx = ky = (T)0;
        return;
    }
    kx = g.kxs * (T)c;
    const int my = (j < g.ny / 2) ? j : j - g.ny;
    ky = g.kys * (T)my;
}

template <typename T>
__device__ __forceinline__ cplx<T> cik(cplx<T> v, T k) { return cplx<T>{-v.y * k, v.x * k}; }

// Divergence of the stresses from their spectra (divstress_uv.f90:21-85,
// divstress_w.f90:21-116) at mode o of level k: divtx, divty on 1..nz-1
// (d/dz of txz, tyz from k to k+1), divtz on 1..nz (d/dz of tzz from k-1 to k
// on 2..nz-1 only); the ky Nyquist row carries no x, y derivative.
template <typename T>
struct StressSpec {
    const cplx<T> *txx, *txy, *tyy, *txz, *tyz, *tzz;
};

template <typename T>
__device__ __forceinline__ cplx<T> divt_xy(const cplx<T>* a, const cplx<T>* b, const cplx<T>* zt,
                                           long o, long S, bool nyq, T kx, T ky, T rdz) {
    const cplx<T> p = a[o], q = b[o];
    const cplx<T> z1 = zt[o + S], z0 = zt[o];
    const cplx<T> dxy = nyq ? cplx<T>{(T)0, (T)0}
                            : cplx<T>{-p.y * kx + -q.y * ky, p.x * kx + q.x * ky};
    return cplx<T>{dxy.x + rdz * (z1.x - z0.x), dxy.y + rdz * (z1.y - z0.y)};
}

template <typename T>
__device__ __forceinline__ cplx<T> divt_z(const StressSpec<T>& t, long o, long S, int k, int nz,
                                          bool nyq, T kx, T ky, T rdz) {
    const cplx<T> p = t.txz[o], q = t.tyz[o];
    cplx<T> z = nyq ? cplx<T>{(T)0, (T)0}
                    : cplx<T>{-p.y * kx + -q.y * ky, p.x * kx + q.x * ky};
    if (k >= 2 && k <= nz - 1) {
        const cplx<T> z1 = t.tzz[o], z0 = t.tzz[o - S];
        z = cplx<T>{z.x + rdz * (z1.x - z0.x), z.y + rdz * (z1.y - z0.y)};
    }
    return z;
}

// RHS assembly and AB2 update (main.f90:211-296) on the spectra, levels
// 1..nz: R = -conv - divt (+ mean pressure force on the (0,0) mode)
// (+ the disks' forcing on levels kf0..kf1) for x, y on 1..nz-1 and z on
// 1..nz (no forcing at nz); members with euler[m] take R_f = R; then
// u* = u + dt (t1 R + t2 R_f) in place of u (x, y on 1..nz-1, z on 1..nz).
// divt is formed here from the stress spectra; its z component at levels 1
// and nz (the pressure solve's boundary rows) is also stored in dvz_bc.
template <typename T>
struct SpecAb2 {
    const cplx<T> *cx, *cy, *cz;       // advection spectra
    StressSpec<T> ts;                  // stress spectra, levels 1..nz
    cplx<T>* dvz_bc;                   // divtz spectra (levels 1 and nz written)
    T rdz;
    const cplx<T> *fx, *fy, *fz;       // forcing spectra (levels kf0..kf1) or null
    cplx<T> *rx, *ry, *rz;             // R out
    cplx<T> *rfx, *rfy, *rfz;          // R_f in (out for Euler members)
    cplx<T> *ux, *uy, *uz;             // u in, u* out
    const double* mpf;                 // 2 per member
    const int* euler;
    int use_mpf, kf0, kf1;
    T dt, t1, t2;
};

template <typename T>
__global__ void k_spec_ab2(SpecAb2<T> a, SpecGeo<T> g) {
    int m, k, j, c;
    long o;
    if (!spec_of(g, m, k, j, c, o)) return;
    const bool eu = a.euler[m] != 0;
    const bool forced = a.fx && k >= a.kf0 && k <= a.kf1 && k <= g.nz - 1;
    const bool zero_mode = (j == 0 && c == 0);
    T kx, ky;
    spec_k(g, j, c, kx, ky);
    const bool nyq = j == g.ny / 2;
    const long S = g.splane;
    auto upd = [&](const cplx<T>* cv, const cplx<T> dd, const cplx<T>* fv, cplx<T>* r,
                   cplx<T>* rf, cplx<T>* u, T mp, bool mpf_on) {
        const cplx<T> cc = cv[o];
        cplx<T> R{-cc.x - dd.x, -cc.y - dd.y};
        if (mpf_on && zero_mode) R.x = R.x + mp;
        if (fv) {
            const cplx<T> f = fv[o];
            R = cplx<T>{R.x + f.x, R.y + f.y};
        }
        r[o] = R;
        cplx<T> F = R;
        if (eu) rf[o] = R;
        else F = rf[o];
        const cplx<T> uu = u[o];
        u[o] = cplx<T>{uu.x + a.dt * (a.t1 * R.x + a.t2 * F.x),
                       uu.y + a.dt * (a.t1 * R.y + a.t2 * F.y)};
    };
    if (k <= g.nz - 1) {
        const cplx<T> dvx = divt_xy(a.ts.txx, a.ts.txy, a.ts.txz, o, S, nyq, kx, ky, a.rdz);
        const cplx<T> dvy = divt_xy(a.ts.txy, a.ts.tyy, a.ts.tyz, o, S, nyq, kx, ky, a.rdz);
        upd(a.cx, dvx, forced ? a.fx : nullptr, a.rx, a.rfx, a.ux, (T)a.mpf[2 * m],
            a.use_mpf != 0);
        upd(a.cy, dvy, forced ? a.fy : nullptr, a.ry, a.rfy, a.uy, (T)a.mpf[2 * m + 1],
            a.use_mpf != 0);
    }
    const cplx<T> dvz = divt_z(a.ts, o, S, k, g.nz, nyq, kx, ky, a.rdz);
    if (k == 1 || k == g.nz) a.dvz_bc[o] = dvz;
    upd(a.cz, dvz, forced ? a.fz : nullptr, a.rz, a.rfz, a.uz, (T)0, false);
}

// RHS -= grad p and the projection (main.f90:321-326, forcing.f90:149-244)
// on the spectra, levels 1..nz: dpdx, dpdy = i kx p, i ky p (1..nz-1);
// dpdz = (p(k) - p(k-1)) / dz (1..nz); u = u* - dt t1 dpdx (1..nz-1),
// w = w* - dt t1 dpdz (2..nz-1).  The lid and wall rows follow in
// k_spec_bc.
template <typename T>
__global__ void k_spec_project(const cplx<T>* p, cplx<T>* rx, cplx<T>* ry, cplx<T>* rz,
                               cplx<T>* ux, cplx<T>* uy, cplx<T>* uz, SpecGeo<T> g, T dt, T t1,
                               T dzv) {
    int m, k, j, c;
    long o;
    if (!spec_of(g, m, k, j, c, o)) return;
    T kx, ky;
    spec_k(g, j, c, kx, ky);
    const cplx<T> pk = p[o], pm = p[o - g.splane];
    const cplx<T> gz{(pk.x - pm.x) / dzv, (pk.y - pm.y) / dzv};
    if (k <= g.nz - 1) {
        const cplx<T> gx = cik(pk, kx), gy = cik(pk, ky);
        cplx<T> r = rx[o];
        rx[o] = cplx<T>{r.x - gx.x, r.y - gx.y};
        r = ry[o];
        ry[o] = cplx<T>{r.x - gy.x, r.y - gy.y};
        cplx<T> u = ux[o];
        ux[o] = cplx<T>{u.x + dt * (-(t1 * gx.x)), u.y + dt * (-(t1 * gx.y))};
        u = uy[o];
        uy[o] = cplx<T>{u.x + dt * (-(t1 * gy.x)), u.y + dt * (-(t1 * gy.y))};
    }
    cplx<T> r = rz[o];
    rz[o] = cplx<T>{r.x - gz.x, r.y - gz.y};
    if (k >= 2 && k <= g.nz - 1) {
        const cplx<T> u = uz[o];
        uz[o] = cplx<T>{u.x + dt * (-(t1 * gz.x)), u.y + dt * (-(t1 * gz.y))};
    }
}

// Boundary rows of the projected velocity (forcing.f90:226-243, main.f90:
// 298-308): u(nz) = u(nz-1), v(nz) = v(nz-1), w(nz) = w(1) = 0, and level 0
// BOGUS (the spectrum of a constant plane: its (0,0) mode).
template <typename T>
__global__ void k_spec_bc(cplx<T>* ux, cplx<T>* uy, cplx<T>* uz, SpecGeo<T> g) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.ny * g.NK) return;
    const long mb = (long)blockIdx.y * g.msp;
    const long S = g.splane;
    const long top = mb + (long)g.nz * S + idx;
    ux[top] = ux[top - S];
    uy[top] = uy[top - S];
    const cplx<T> zero{(T)0, (T)0};
    uz[top] = zero;
    uz[mb + S + idx] = zero;
    const cplx<T> b = idx == 0 ? cplx<T>{(T)BOGUS, (T)0} : zero;
    ux[mb + idx] = b;
    uy[mb + idx] = b;
    uz[mb + idx] = b;
}

// dst += dt (t1 hi + t2 (sel[m] ? hi : hp)) on levels of the geometry
// (the high-wavenumber remainders of the forcing that the spectra drop; a
// null hi or hp counts as zero).
template <typename T>
__global__ void k_add_high(T* dst, const T* hi, const T* hp, const int* sel, Geo g, T dt, T t1,
                           T t2) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.ld, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    const T h = hi ? hi[o] : (T)0;
    const T b = (sel && sel[m]) ? h : (hp ? hp[o] : (T)0);
    dst[o] = dst[o] + dt * (t1 * h + t2 * b);
}

// g = (f ? hp(f) : 0) - d/dz hp(t) (Case::rhs_high), levels 1..nz-1 of the
// geometry: x uses hp(txz)(k+1) - hp(txz)(k), y hp(tyz), z hp(tzz)(k) -
// hp(tzz)(k-1) on 2..nz-1 only.  g holds hp(f) on entry when has_f.
template <typename T>
__global__ void k_rhs_high(T* gx, T* gy, T* gz, const T* hxz, const T* hyz, const T* hzz,
                           int has_f, Geo g, T rdz) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.ld, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    const long P = g.plane;
    const T fx = has_f ? gx[o] : (T)0, fy = has_f ? gy[o] : (T)0, fz = has_f ? gz[o] : (T)0;
    gx[o] = fx - rdz * (hxz[o + P] - hxz[o]);
    gy[o] = fy - rdz * (hyz[o + P] - hyz[o]);
    gz[o] = (k >= 2 && k <= g.nz - 1) ? fz - rdz * (hzz[o] - hzz[o - P]) : fz;
}

// The spectra of members starting from physical fields replace their state:
// dst[q] = src[q] for members with sel[m] (grid.y = member, grid.z = q).
template <typename T>
struct SelArgs {
    const cplx<T>* src[6];
    cplx<T>* dst[6];
};

template <typename T>
__global__ void k_select_members(SelArgs<T> a, const int* sel, long msp, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (i >= n || !sel[m]) return;
    const long o = (long)m * msp + i;
    a.dst[blockIdx.z][o] = a.src[blockIdx.z][o];
}

// dst += (sel[m] ? hi : hp) (RHS_f of an Euler member is this step's RHS)
template <typename T>
__global__ void k_add_sel(T* dst, const T* hi, const T* hp, const int* sel, Geo g) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.ld, m, k, j, i, mb)) return;
    if (!sel[m]) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    dst[o] = dst[o] + hi[o] - (hp ? hp[o] : (T)0);
}

// dst = a - b (the high part of a field: the field minus its band-limited
// round trip), levels of the geometry.
template <typename T>
__global__ void k_sub(T* dst, const T* a, const T* b, Geo g) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.ld, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    dst[o] = a[o] - b[o];
}

// dst += src (levels of the geometry)
template <typename T>
__global__ void k_add(T* dst, const T* src, Geo g) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.ld, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    dst[o] = dst[o] + src[o];
}

}  // namespace lg
