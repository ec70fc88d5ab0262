// Pressure Poisson solve in (kx, ky) space: RHS assembly, batched Thomas
// and the (0,0) mode (press_stag_array.f90:149-250, tridag_array.f90:165-246).
// Always computed in FP64, whatever the storage type of the spectra.
#pragma once

#include "common.hpp"

namespace lg {

struct TriGeo {
    int ny, NK, nz;
    long splane;         // ny * NK complex per level
    long smstride;       // complex elements between members
    int M;
    double kxs, kys;     // 2 pi / Lx, 2 pi / Ly
    double dz, const3, const4;
    double hsc = 1.0;    // factor on Hx, Hy, Hz (1/(tadv1 dt) for velocity spectra)
};

// The Thomas elimination coefficients gam(m) = c(m-1) / bet(m-1) of every
// (ky, kx) system: they depend on the grid only (kx, ky and dz), not on the
// member or the step, so they are formed once per case, by the recurrence
// k_tridag's forward sweep runs (the same operations, so the same values),
// into gam[(nz+1) * ny * NK] (FP64).
static __global__ void __launch_bounds__(256) k_tridag_gam(double* gam, TriGeo g) {
    // no FMA contraction here or in k_tridag: both form bet = b - a gam with
    // two roundings (as the reference does), so the table holds the values
    // the forward sweep uses
#pragma clang fp contract(off)
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.ny * g.NK) return;
    const int jy = idx / g.NK;
    const int jx = idx - jy * g.NK;
    const long S = g.splane;
    const double kx = g.kxs * (double)jx;
    const int my = (jy < g.ny / 2) ? jy : jy - g.ny;
    const double ky = g.kys * (double)my;
    const double a_in = g.const3, c_in = g.const3;
    const double b_in = -(kx * kx + ky * ky + 2.0 * g.const3);
    double bet = -1.0, cprev = 1.0;
    gam[idx] = 0.0;
    for (int m = 1; m <= g.nz; ++m) {
        const double a = m < g.nz ? a_in : -1.0;
        const double b = m < g.nz ? b_in : 1.0;
        const double cc = m < g.nz ? c_in : 0.0;
        const double gm = cprev / bet;
        gam[(long)m * S + idx] = gm;
        bet = b - a * gm;
        cprev = cc;
    }
}

// One thread per (member, ky, kx < nx/2).  Hx, Hy hold levels 1..nz-1, Hz levels
// 1..nz (all already scaled by 1/(nx ny tadv1 dt)); rb, rt the bottom/top
// Neumann planes scaled by 1/(nx ny).  Writes p_hat for rows 0..nz and
// uses usc (FP64 complex, (nz+1) * ny * NK per member) as the Thomas
// scratch; the forward sweep forms gam in registers and the back
// substitution reads it from the case's table (k_tridag_gam, shared by the
// members).
// One column per thread leaves few waves per SIMD (ny * NK / 64 in all), so
// each sweep issues the loads of a block of TRI_BLK levels together and then
// runs the recurrence through them: one memory round trip per block, not
// per level.  The arithmetic is the plain per-level Thomas recurrence.
template <typename T>
struct TriBlk {
    static constexpr int n = sizeof(T) == 4 ? 16 : 8;
};

template <typename T>
__global__ void __launch_bounds__(256) k_tridag(const cplx<T>* Hx, const cplx<T>* Hy, const cplx<T>* Hz,
                         const cplx<T>* rb, const cplx<T>* rt, cplx<T>* ph, const double* gam,
                         double2* usc, TriGeo g, int* err) {
#pragma clang fp contract(off)
    constexpr int B = TriBlk<T>::n;
    const int gidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (gidx >= g.M * g.ny * g.NK) return;
    const int mi = gidx / (g.ny * g.NK);
    const int idx = gidx - mi * (g.ny * g.NK);
    const long mb = (long)mi * g.smstride;
    Hx += mb;
    Hy += mb;
    Hz += mb;
    rb += mb;
    rt += mb;
    ph += mb;
    usc += mb;
    const int jy = idx / g.NK;
    const int jx = idx - jy * g.NK;
    const long S = g.splane;
    const int n = g.nz + 1;
    auto ld = [](const cplx<T>* p, long o) -> double2 {
        const cplx<T> v = p[o];
        return make_double2((double)v.x, (double)v.y);
    };
    const double hs = g.hsc;
    auto sc = [&](cplx<T> v) -> double2 { return make_double2((double)v.x * hs, (double)v.y * hs); };
    auto st = [&](long o, double2 v) { ph[o] = cplx<T>{(T)v.x, (T)v.y}; };
    if (jy == g.ny / 2) {
        for (int m = 0; m < n; ++m) st((long)m * S + idx, make_double2(0.0, 0.0));
        return;
    }
    if (jx == 0 && jy == 0) {
        // zero-wavenumber mode: p(0) = 0, p(1) = -dz rb, p(k) = p(k-1) + Hz(k) dz
        double2 p0 = make_double2(0.0, 0.0);
        st(idx, p0);
        const double2 b = ld(rb, idx);
        double2 pk = make_double2(p0.x - g.dz * b.x, p0.y - g.dz * b.y);
        st(S + idx, pk);
        for (int m0 = 2; m0 <= g.nz; m0 += B) {
            cplx<T> hz[B];
#pragma unroll
            for (int i = 0; i < B; ++i) hz[i] = Hz[(long)min(m0 + i, g.nz) * S + idx];
#pragma unroll
            for (int i = 0; i < B; ++i) {
                const int m = m0 + i;
                if (m <= g.nz) {
                    const double2 h = sc(hz[i]);
                    pk = make_double2(pk.x + h.x * g.dz, pk.y + h.y * g.dz);
                    st((long)m * S + idx, pk);
                }
            }
        }
        return;
    }
    const double kx = g.kxs * (double)jx;
    const int my = (jy < g.ny / 2) ? jy : jy - g.ny;
    const double ky = g.kys * (double)my;
    const double a_in = g.const3, c_in = g.const3;
    const double b_in = -(kx * kx + ky * ky + 2.0 * g.const3);
    // row 0: b = -1, c = 1, r = -dz rb
    double2 r0 = ld(rb, idx);
    double2 u = make_double2((-g.dz * r0.x) / -1.0, (-g.dz * r0.y) / -1.0);
    usc[idx] = u;
    double bet = -1.0;
    double cprev = 1.0;
    // Hz(m) of the level being solved (Hz(1) before the first)
    double2 z0 = sc(Hz[S + idx]);
    for (int m0 = 1; m0 <= g.nz; m0 += B) {
        // rows m < nz read Hx(m), Hy(m), Hz(m + 1); clamped indices keep the
        // block's loads in bounds past its end
        cplx<T> hx[B], hy[B], hz[B];
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const long o = (long)min(m0 + i, g.nz - 1) * S + idx;
            hx[i] = Hx[o];
            hy[i] = Hy[o];
            hz[i] = Hz[o + S];
        }
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const int m = m0 + i;
            if (m > g.nz) continue;
            double a, b, cc;
            double2 r;
            if (m < g.nz) {
                a = a_in;
                b = b_in;
                cc = c_in;
                const double2 vx = sc(hx[i]);
                const double2 vy = sc(hy[i]);
                const double2 z1 = sc(hz[i]);
                // i kx Hx + i ky Hy + (Hz(m+1) - Hz(m)) / dz
                const double ahx_r = -vx.y * kx, ahx_i = vx.x * kx;
                const double ahy_r = -vy.y * ky, ahy_i = vy.x * ky;
                r.x = (ahx_r + ahy_r) + (z1.x - z0.x) * g.const4;
                r.y = (ahx_i + ahy_i) + (z1.y - z0.y) * g.const4;
                z0 = z1;
            } else {
                a = -1.0;
                b = 1.0;
                cc = 0.0;
                const double2 t = ld(rt, idx);
                r = make_double2(-g.dz * t.x, -g.dz * t.y);
            }
            const double gm = cprev / bet;
            bet = b - a * gm;
            if (bet == 0.0) atomicExch(err, 1);
            u = make_double2((r.x - a * u.x) / bet, (r.y - a * u.y) / bet);
            usc[(long)m * S + idx] = u;
            cprev = cc;
        }
    }
    st((long)g.nz * S + idx, u);
    double2 up = u;
    // back substitution, levels nz-1 .. 0 in blocks
    for (int m0 = g.nz - 1; m0 >= 0; m0 -= B) {
        double gb[B];
        double2 ub[B];
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const int m = max(m0 - i, 0);
            gb[i] = gam[(long)(m + 1) * S + idx];
            ub[i] = usc[(long)m * S + idx];
        }
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const int m = m0 - i;
            if (m < 0) continue;
            const double gm = gb[i];
            double2 um = ub[i];
            um = make_double2(um.x - gm * up.x, um.y - gm * up.y);
            st((long)m * S + idx, um);
            up = um;
        }
    }
}

}  // namespace lg
