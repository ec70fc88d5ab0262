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

// Divergence of the stresses from their spectra (levels 1..nz of txx, txy,
// tyy, txz, tyz, tzz): divtx, divty on 1..nz-1 (d/dz of txz, tyz from k to
// k+1), divtz on 1..nz (d/dz of tzz from k-1 to k on 2..nz-1 only).
template <typename T>
__global__ void k_spec_divt(const cplx<T>* txx, const cplx<T>* txy, const cplx<T>* tyy,
                            const cplx<T>* txz, const cplx<T>* tyz, const cplx<T>* tzz,
                            cplx<T>* dvx, cplx<T>* dvy, cplx<T>* dvz, SpecGeo<T> g, T rdz) {
    int m, k, j, c;
    long o;
    if (!spec_of(g, m, k, j, c, o)) return;
    T kx, ky;
    spec_k(g, j, c, kx, ky);
    const long S = g.splane;
    const bool nyq = j == g.ny / 2;
    if (k <= g.nz - 1) {
        const cplx<T> a = txx[o], b = txy[o], e = tyy[o];
        const cplx<T> xz1 = txz[o + S], xz0 = txz[o], yz1 = tyz[o + S], yz0 = tyz[o];
        const cplx<T> dxy = nyq ? cplx<T>{(T)0, (T)0}
                                : cplx<T>{-a.y * kx + -b.y * ky, a.x * kx + b.x * ky};
        const cplx<T> dyy = nyq ? cplx<T>{(T)0, (T)0}
                                : cplx<T>{-b.y * kx + -e.y * ky, b.x * kx + e.x * ky};
        dvx[o] = cplx<T>{dxy.x + rdz * (xz1.x - xz0.x), dxy.y + rdz * (xz1.y - xz0.y)};
        dvy[o] = cplx<T>{dyy.x + rdz * (yz1.x - yz0.x), dyy.y + rdz * (yz1.y - yz0.y)};
    } else {
        dvx[o] = cplx<T>{(T)0, (T)0};
        dvy[o] = cplx<T>{(T)0, (T)0};
    }
    const cplx<T> p = txz[o], q = tyz[o];
    cplx<T> z = nyq ? cplx<T>{(T)0, (T)0}
                    : cplx<T>{-p.y * kx + -q.y * ky, p.x * kx + q.x * ky};
    if (k >= 2 && k <= g.nz - 1) {
        const cplx<T> z1 = tzz[o], z0 = tzz[o - S];
        z = cplx<T>{z.x + rdz * (z1.x - z0.x), z.y + rdz * (z1.y - z0.y)};
    }
    dvz[o] = z;
}

// RHS assembly and AB2 update (main.f90:211-296) on the spectra, levels
// 1..nz: R = -conv - divt (+ mean pressure force on the (0,0) mode)
// (+ the disks' forcing on levels kf0..kf1) for x, y on 1..nz-1 and z on
// 1..nz (no forcing at nz); members with euler[m] take R_f = R; then
// u* = u + dt (t1 R + t2 R_f) in place of u (x, y on 1..nz-1, z on 1..nz).
template <typename T>
struct SpecAb2 {
    const cplx<T> *cx, *cy, *cz;       // advection spectra
    const cplx<T> *dx, *dy, *dz;       // stress divergence spectra
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
    auto upd = [&](const cplx<T>* cv, const cplx<T>* dv, const cplx<T>* fv, cplx<T>* r,
                   cplx<T>* rf, cplx<T>* u, T mp, bool mpf_on) {
        const cplx<T> cc = cv[o], dd = dv[o];
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
        upd(a.cx, a.dx, forced ? a.fx : nullptr, a.rx, a.rfx, a.ux, (T)a.mpf[2 * m],
            a.use_mpf != 0);
        upd(a.cy, a.dy, forced ? a.fy : nullptr, a.ry, a.rfy, a.uy, (T)a.mpf[2 * m + 1],
            a.use_mpf != 0);
    }
    upd(a.cz, a.dz, forced ? a.fz : nullptr, a.rz, a.rfz, a.uz, (T)0, false);
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

// divt += the stress high parts' vertical derivatives (levels 1..nz-1):
// the physical divtx, divty, divtz of the reference (divstress_uv.f90,
// divstress_w.f90) from their spectra.
template <typename T>
__global__ void k_divt_high(T* dx, T* dy, T* dz, const T* hxz, const T* hyz, const T* hzz,
                            Geo g, T rdz) {
    int m, k, j, i;
    long mb;
    if (!pt_of(g, g.ld, m, k, j, i, mb)) return;
    const long o = mb + (long)k * g.plane + (long)j * g.ld + i;
    const long P = g.plane;
    dx[o] = dx[o] + rdz * (hxz[o + P] - hxz[o]);
    dy[o] = dy[o] + rdz * (hyz[o + P] - hyz[o]);
    if (k >= 2) dz[o] = dz[o] + rdz * (hzz[o] - hzz[o - P]);
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
