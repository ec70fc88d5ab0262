"""Pins the oracle's per-subsystem restatements with closed forms, a dense
linear solve, exact-interpolation and invariance properties."""
import math

import numpy as np
import pytest

from oracle import lesgo_oracle as o


def small(nx=16, ny=12, Nz=10, **kw):
    p = o.Params(nx, ny, Nz, **kw)
    return p, o.Spectral(p), o.State.zeros(p)


def test_ddz_linear():
    p, sp, s = small()
    z = (np.arange(p.nz + 1) - 0.5) * p.dz
    f = np.broadcast_to((3.0 * z + 1.0)[:, None, None], (p.nz + 1, p.ny, p.ld)).copy()
    d = o.ddz_uv(p, f)
    np.testing.assert_allclose(d[2:p.nz, :, :p.nx], 3.0, rtol=1e-12)
    assert (d[1] == o.BOGUS).all() and (d[p.nz] == o.BOGUS).all()
    d = o.ddz_w(p, f)
    np.testing.assert_allclose(d[1:p.nz, :, :p.nx], 3.0, rtol=1e-12)
    assert (d[0] == o.BOGUS).all() and (d[p.nz] == o.BOGUS).all()


def test_wallstress_loglaw():
    """Uniform plane (U, V): test filter keeps the mean, so
    u* = kappa |U| / ln(0.5 dz / z0), txz = -u*^2 U/|U|, dudz = u*/(0.5 dz kappa) U/|U|."""
    p, sp, s = small()
    U, V = 7.0, -2.0
    s.u[1, :, :p.nx] = U
    s.v[1, :, :p.nx] = V
    o.wallstress(sp, s)
    M = math.hypot(U, V)
    us = 0.4 * M / math.log(0.5 * p.dz / p.zo)
    np.testing.assert_allclose(s.txz[1, :, :p.nx], -us * us * U / M, rtol=1e-12)
    np.testing.assert_allclose(s.tyz[1, :, :p.nx], -us * us * V / M, rtol=1e-12)
    np.testing.assert_allclose(s.dudz[1, :, :p.nx], us / (0.5 * p.dz * 0.4) * U / M, rtol=1e-12)
    np.testing.assert_allclose(s.dvdz[1, :, :p.nx], us / (0.5 * p.dz * 0.4) * V / M, rtol=1e-12)
    assert (s.txz[p.nz] == 0).all() and (s.dudz[p.nz] == 0).all()


def test_sij_and_sgs_linear_shear():
    """u = a z (so dudz = a at w nodes), constant Cs: Nu_t = Cs D^2 |S|,
    |S| = a, S13 = a/2, txz = -2 Nu_t S13 on interior w nodes."""
    p, sp, s = small()
    a, cs = 2.5, 0.03
    s.dudz[:, :, :p.nx] = a
    s.Cs_opt2[:] = cs
    s.jt_total = 1            # not a dynamic step
    o.sgs_stag(sp, s)
    K = np.s_[2:p.nz - 1, :, :p.nx]
    np.testing.assert_allclose(s.S13[K], a / 2, rtol=1e-14)
    nut = cs * p.delta ** 2 * a
    np.testing.assert_allclose(s.Nu_t[K], nut, rtol=1e-12)
    np.testing.assert_allclose(s.txz[K], -2 * nut * a / 2, rtol=1e-12)
    np.testing.assert_allclose(s.txx[K], 0.0, atol=1e-15)


def test_tridag_dense():
    """Batched Thomas equals a dense solve of each (kx, ky) system."""
    p, sp, s = small(8, 6, 5)
    rng = np.random.default_rng(3)
    n = p.nz + 1
    a = rng.uniform(0.5, 1, (n, p.ny, p.lh))
    c = rng.uniform(0.5, 1, (n, p.ny, p.lh))
    b = -(a + c + rng.uniform(1, 2, (n, p.ny, p.lh)))
    r = rng.standard_normal((n, p.ny, p.lh)) + 1j * rng.standard_normal((n, p.ny, p.lh))
    u = o.tridag_array(p, a, b, c, r)
    for ky in range(p.ny):
        for kx in range(p.lh - 1):
            if ky == p.ny // 2 or (kx == 0 and ky == 0):
                continue
            A = np.zeros((n, n))
            for m in range(n):
                A[m, m] = b[m, ky, kx]
                if m > 0:
                    A[m, m - 1] = a[m, ky, kx]
                if m < n - 1:
                    A[m, m + 1] = c[m, ky, kx]
            np.testing.assert_allclose(u[:, ky, kx], np.linalg.solve(A, r[:, ky, kx]), rtol=1e-11, atol=1e-12)


def spectral_divergence(sp, s):
    """Same operator as rmsdiv: spectral d/dx, d/dy + FD d/dz on uv levels."""
    p = sp.p
    dudx = o.ddx(sp, s.u)
    dvdy = o.ddy(sp, s.v)
    K = np.s_[1:p.nz, :, :p.nx]
    dwdz = (s.w[2:p.nz + 1, :, :p.nx] - s.w[1:p.nz, :, :p.nx]) / p.dz
    return dudx[K] + dvdy[K] + dwdz


def random_ustar(p, sp, seed=4):
    rng = np.random.default_rng(seed)
    s = o.State.zeros(p)
    for n in ("u", "v", "w"):
        f = rng.standard_normal((p.nz + 1, p.ny, p.ld))
        s.f[n][:] = o.filt_da(sp, f)[0]          # band limit (no Nyquist)
    s.w[1] = 0.0
    s.w[p.nz] = 0.0
    return s


def test_projection_divergence_free_and_idempotent():
    p, sp, _ = small(16, 12, 10)
    s = random_ustar(p, sp)
    s.divtz[:] = 0.0
    o.press_stag_array(sp, s)
    o.project(p, s)
    div = spectral_divergence(sp, s)
    assert np.abs(div).max() < 1e-9 * np.abs(s.u).max()
    # projecting a divergence-free field again changes it only at roundoff
    u0, v0, w0 = s.u.copy(), s.v.copy(), s.w.copy()
    o.press_stag_array(sp, s)
    o.project(p, s)
    X = np.s_[1:p.nz, :, :p.nx]
    assert np.abs(s.u[X] - u0[X]).max() < 1e-10
    assert np.abs(s.w[X] - w0[X]).max() < 1e-10


def test_pressure_neumann_rows():
    """The bottom/top rows impose (p(1)-p(0))/dz = -divtz(1) plane by plane."""
    p, sp, _ = small(8, 8, 6)
    s = random_ustar(p, sp, 5)
    rng = np.random.default_rng(6)
    s.divtz[1, :, :p.nx] = o.filt_da(sp, rng.standard_normal((p.ny, p.ld)))[0][:, :p.nx]
    o.press_stag_array(sp, s)
    dpdz0 = (s.f["p"][1, :, :p.nx] - s.f["p"][0, :, :p.nx]) / p.dz
    np.testing.assert_allclose(dpdz0, -s.divtz[1, :, :p.nx], atol=1e-10)


def test_divstress_trig():
    from test_oracle_fft import trig_field
    p, sp, s = small(16, 12, 8)
    f, fx, fy = trig_field(p, [(2, 1, 1.0), (3, -4, 0.5)])
    g, gx, gy = trig_field(p, [(1, 2, 0.7)], shift=1.1)
    for k in range(p.nz + 1):
        s.txx[k] = f * (k + 1)
        s.txy[k] = g
        s.tyy[k] = 2 * g
        s.txz[k] = f * k
        s.tyz[k] = g * k
        s.tzz[k] = f * k * k
    o.divstress_uv(sp, s)
    o.divstress_w(sp, s)
    for k in range(2, p.nz - 1):
        X = np.s_[:, :p.nx]
        np.testing.assert_allclose(s.divtx[k][X], (fx * (k + 1) + gy + f / p.dz)[X], atol=1e-10)
        np.testing.assert_allclose(s.divty[k][X], (gx + 2 * gy + g / p.dz)[X], atol=1e-10)
        np.testing.assert_allclose(s.divtz[k][X], (fx * k + gy * k + f * (2 * k - 1) / p.dz)[X], atol=1e-9)


def test_convec_band_limited_matches_unpadded_products():
    """With u, v, w and the derivative fields limited to |k| < n/4, every
    product is resolved on the unpadded grid, so the 3/2-rule convec must
    equal the same staggered formulas evaluated pointwise on the nx x ny grid."""
    from test_oracle_fft import trig_field
    p, sp, s = small(16, 16, 6)
    rng = np.random.default_rng(7)
    names = ("u", "v", "w", "dudy", "dudz", "dvdx", "dvdz", "dwdx", "dwdy")
    for n in names:
        for k in range(p.nz + 1):
            modes = [(int(rng.integers(0, 4)), int(rng.integers(-3, 4)), rng.standard_normal()) for _ in range(3)]
            s.f[n][k] = trig_field(p, modes, shift=rng.uniform(0, 6))[0]
    o.convec(sp, s)
    X = np.s_[:, :p.nx]
    u, v, w = s.u, s.v, s.w
    om1 = s.dwdy - s.dvdz
    om2 = s.dudz - s.dwdx
    om3 = s.dvdx - s.dudy
    om1[1] = 0.5 * (s.dwdy[1] + s.dwdy[2]) - s.dvdz[1]
    om2[1] = s.dudz[1] - 0.5 * (s.dwdx[1] + s.dwdx[2])
    nz = p.nz
    for k in range(1, nz):
        if k == 1:
            ex = v[1] * (-om3[1]) + 0.5 * w[2] * om2[2]
            ey = u[1] * om3[1] + 0.5 * w[2] * (-om1[2])
        elif k == nz - 1:
            ex = v[k] * (-om3[k]) + 0.5 * w[k] * om2[k]
            ey = u[k] * om3[k] + 0.5 * w[k] * (-om1[k])
        else:
            ex = v[k] * (-om3[k]) + 0.5 * (w[k + 1] * om2[k + 1] + w[k] * om2[k])
            ey = u[k] * om3[k] + 0.5 * (w[k + 1] * (-om1[k + 1]) + w[k] * (-om1[k]))
        np.testing.assert_allclose(s.RHSx[k][X], ex[X], atol=1e-11)
        np.testing.assert_allclose(s.RHSy[k][X], ey[X], atol=1e-11)
        ez = 0.0 if k == 1 else 0.5 * ((u[k] + u[k - 1]) * (-om2[k]) + (v[k] + v[k - 1]) * om1[k])
        np.testing.assert_allclose(s.RHSz[k][X], (ez + 0 * u[k])[X], atol=1e-11)


def test_interp_linear_exact():
    """Trilinear interpolation reproduces a field linear in x, y, z (away
    from the periodic seam and the wall/lid special cells)."""
    p, sp, s = small(16, 12, 10)
    zw = (np.arange(p.nz + 1) - 1.0) * p.dz
    x = np.arange(p.nx) * p.dx
    y = np.arange(p.ny) * p.dy
    Z, Y, X = np.meshgrid(zw, y, x, indexing="ij")
    var = 1.0 + 0.3 * X - 0.7 * Y + 2.0 * Z
    rng = np.random.default_rng(8)
    px = rng.uniform(0.1, (p.nx - 1.5) * p.dx, 200)
    py = rng.uniform(0.1, (p.ny - 1.5) * p.dy, 200)
    pz = rng.uniform(1.1 * p.dz, (p.nz - 1.5) * p.dz, 200)
    got = o.trilinear_interp_w(p, var, px, py, pz)
    np.testing.assert_allclose(got, 1.0 + 0.3 * px - 0.7 * py + 2.0 * pz, rtol=1e-12)


def test_interpolag_advects_by_one_cell():
    """Uniform u with u*lagran_dt = dx and w = 0 moves F exactly one cell
    downstream in x (interpolag_Sdep.f90:154-176) on interior levels."""
    p, sp, s = small(16, 12, 10)
    U = p.dx / p.lagran_dt
    s.u[:] = U
    rng = np.random.default_rng(9)
    F0 = rng.uniform(0.1, 1.0, (p.nz + 1, p.ny, p.ld))
    s.F_LM[:] = F0
    o.interpolag_Sdep(p, s)
    K = np.s_[2:p.nz]
    np.testing.assert_allclose(s.F_LM[K][:, :, :p.nx], np.roll(F0[K][:, :, :p.nx], 1, axis=2), rtol=1e-9)


def test_lagrange_translation_invariant():
    """Shifting the velocity field by whole cells in x and y shifts Cs_opt2."""
    p, sp, _ = small(16, 12, 8)
    s = random_ustar(p, sp, 10)
    s.u[:] += 5.0
    for n in ("u", "v", "w"):
        s.f[n][1:], a, b = o.filt_da(sp, s.f[n])[0][1:], None, None
    s.jt_total = 4
    base = s.copy()
    o.step(sp, s)
    t = base.copy()
    for n in list(t.f):
        t.f[n][:, :, :p.nx] = np.roll(np.roll(t.f[n][:, :, :p.nx], 3, axis=2), 2, axis=1)
    o.step(sp, t)
    for n in ("Cs_opt2", "F_LM", "F_MM", "u", "w"):
        a = np.roll(np.roll(s.f[n][1:p.nz, :, :p.nx], 3, axis=2), 2, axis=1)
        np.testing.assert_allclose(t.f[n][1:p.nz, :, :p.nx], a, rtol=1e-9, atol=1e-12)


def test_config0_step_is_divergence_free():
    p = o.Params(32, 32, 32)
    s, snaps = o.run(p, 12, seed=0, every=5)
    assert snaps[10][4] < 1e-12
    assert np.isfinite(s.u[1:p.nz]).all()


def test_lagrange_floor_sensitivity_is_intrinsic():
    """Documents the conditioning of the reference algorithm that bounds
    achievable parity in the actuator-disk case: a 1e-14 relative
    perturbation of u at step 5 stays ~1e-14 until the next dynamic update
    (step 10), where Tn = 1.5 delta max(F_LM F_MM, 1e-24)^(-1/8) near the
    clip floor amplifies it by orders of magnitude."""
    p = o.Params(48, 32, 24)
    specs = o.aligned_farm(p, 2, 2, dia=0.3, hub=0.3, sx_d=7, sy_d=5)
    sp = o.Spectral(p)
    s = o.initial_condition(p, 0)
    s.turbines = o.Turbines(p, specs)
    for _ in range(5):
        o.step(sp, s)
    t = s.copy()
    rng = np.random.default_rng(0)
    t.u[1:p.nz] *= 1 + 1e-14 * rng.standard_normal(t.u[1:p.nz].shape)
    X = np.s_[1:p.nz, :, :p.nx]
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    for _ in range(4):
        o.step(sp, s)
        o.step(sp, t)
    assert rel(t.u[X], s.u[X]) < 1e-13
    o.step(sp, s)
    o.step(sp, t)
    assert rel(t.u[X], s.u[X]) > 1e-11


def test_turbine_center_cells_and_sample():
    """Disk centre cells (turbines.f90:196-198): Fortran nint (half away from
    zero) of xloc/dx, yloc/dy, height/dz + 0.5, 1-based; 0-based x, y here,
    a centre left of the first cell taken periodically; the centre sample is
    u, v, w_uv there (turbines.f90:538-547)."""
    p = o.Params(32, 16, 16)
    dx, dy, dz = p.dx, p.dy, p.dz
    specs = [o.TurbineSpec(3.3 * dx, 5.6 * dy, 2.2 * dz, 2 * dy, 0.0),      # (3, 6, 3)
             o.TurbineSpec(0.4 * dx, 7.49 * dy, 4.0 * dz, 2 * dy, 0.0)]     # (0 -> 32, 7, 5): 4.5 -> 5
    t = o.Turbines(p, specs, N=256)
    assert t.centers == [(2, 5, 3), (31, 6, 5)]
    s = o.initial_condition(p, 0)
    s.u[...] = np.arange(s.u.size, dtype=float).reshape(s.u.shape)
    s.v[...] = 2.0
    s.w[...] = np.arange(p.nz + 1, dtype=float)[:, None, None]
    t.forcing(s)
    np.testing.assert_array_equal(t.center_vel[0], [s.u[3, 5, 2], 2.0, 3.5])
    np.testing.assert_array_equal(t.center_vel[1], [s.u[5, 6, 31], 2.0, 5.5])
