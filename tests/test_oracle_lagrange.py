"""Closed-form pins of the oracle's Lagrangian scale-dependent dynamic model
(lagrange_Sdep.f90:88-408) and its near-wall interpolation cells
(functions.f90:413-452).

The velocity and strain fields are single Fourier modes in x chosen against
the sharp test filters (G = 1/(nx ny) for k^2 < (pi/Delta_t)^2, Delta_t =
alpha sqrt(dx dy), alpha = 2 and 4; test_filtermodule.f90:56-66): on the
32 x 32 grid (Lx = 2 pi) the x mode m = 7 passes the 2-delta filter, is
removed by the 4-delta filter, and its square (m = 14) is removed by both.
Then every filtered product is known by hand (trigonometric identities,
no transform):

  u' = Au cos t, v' = Av sin t, w' = Aw cos t   (t = 7 x, on top of means)
  L_ij = filt(u_i u_j) - filt(u_i) filt(u_j) = - (oscillating part of u_i' u_j')
  Q_ij = tfilt(u_i u_j) - tfilt(u_i) tfilt(u_j) = mean of u_i' u_j'
  S11 = As cos t, S22 = -As cos t, S12 = As sin t, S13 = s3 (so |S| = 2 sqrt(As^2 + s3^2)
  is uniform): S_bar = S, S_hat = (0, 0, s3, 0, 0, 0),
  M_ij = 2 Delta^2 (|S| S_ij - 4 |S| S_ij), N_ij = 2 Delta^2 (|S| S_hat_ij - 16 |S_hat| S_hat_ij).

The amplitudes vary linearly with height, so the staggering of u_bar, v_bar,
w_bar is pinned too: on interior w-nodes the average of the two uv levels is
the value at the w-node; level 1 takes u(1) and 0.25 w(2), the lid level
u(nz-1) and w = 0 (lagrange_Sdep.f90:93-117).

The running averages are pinned through their time weights: F_* = 1e100
(Tn = 1.5 Delta 1e-25, clipped up to 1e-24, so epsi rounds to exactly 1 and F_* becomes the
instantaneous LM, MM, QN, NN), F_* = 16 with lagran_dt = 0.75 Delta (Tn =
1.5 Delta 256^(-1/8) = lagran_dt, epsi = 1/2), and the fresh start's first
call with lagran_dt = 0 (F_MM = MM, F_LM = 0.03 MM, epsi = 0)."""
import math

import numpy as np
import pytest

from oracle import lesgo_oracle as o

MODE = 7
U0, V0 = 2.0, -0.5
AS, S3 = 1.0, 0.5


def amp_u(z):
    return 0.6 + 0.8 * z


def amp_v(z):
    return 0.4 - 0.3 * z


def amp_w(z):
    return -(0.03 + 0.05 * z)


def setup(initu=True):
    p = o.Params(32, 32, 8, initu=initu)
    sp = o.Spectral(p)
    s = o.State.zeros(p)
    # the chosen mode sits where the filters are meant to put it
    k = MODE * 2 * math.pi / p.lx
    kc2 = (math.pi / (2 * math.sqrt(p.dx * p.dy))) ** 2
    kc4 = (math.pi / (4 * math.sqrt(p.dx * p.dy))) ** 2
    assert k * k < kc2 and k * k >= kc4 and (2 * k) ** 2 >= kc2 and 2 * MODE < p.nx // 2
    x = np.arange(p.nx) * p.dx
    t = k * x
    nz, dz = p.nz, p.dz
    zuv = lambda kk: (kk - 0.5) * dz
    zw = lambda kk: (kk - 1.0) * dz
    for kk in range(1, nz):
        s.u[kk, :, :p.nx] = U0 + amp_u(zuv(kk)) * np.cos(t)
        s.v[kk, :, :p.nx] = V0 + amp_v(zuv(kk)) * np.sin(t)
    s.u[nz] = s.u[nz - 1]
    s.v[nz] = s.v[nz - 1]
    for kk in range(2, nz):
        s.w[kk, :, :p.nx] = amp_w(zw(kk)) * np.cos(t)
    for kk in range(1, nz + 1):
        s.S11[kk, :, :p.nx] = AS * np.cos(t)
        s.S22[kk, :, :p.nx] = -AS * np.cos(t)
        s.S12[kk, :, :p.nx] = AS * np.sin(t)
        s.S13[kk, :, :p.nx] = S3
    s.jt = s.jt_total = p.cs_count
    return p, sp, s, t


def expected(p, t):
    """LM, MM, QN, NN on levels 1..nz (arrays [nz, nx]) from the
    trigonometric identities above."""
    nz, dz = p.nz, p.dz
    d2 = p.delta ** 2
    sigma = 2.0 * math.sqrt(AS ** 2 + S3 ** 2)
    c, s_, c2, s2 = np.cos(t), np.sin(t), np.cos(2 * t), np.sin(2 * t)
    LM, MM, QN, NN = (np.empty((nz, p.nx)) for _ in range(4))
    for kk in range(1, nz + 1):
        if kk == 1:                       # uv level 1 (lagrange_Sdep.f90:99-101)
            au, av, aw = amp_u(0.5 * dz), amp_v(0.5 * dz), 0.25 * amp_w(dz)
        elif kk == nz:                    # stress-free lid (104-107)
            au, av, aw = amp_u((nz - 1.5) * dz), amp_v((nz - 1.5) * dz), 0.0
        else:                             # the w node itself (113-116)
            zwk = (kk - 1.0) * dz
            au, av, aw = amp_u(zwk), amp_v(zwk), amp_w(zwk)
        L11, L22, L33 = -0.5 * au * au * c2, 0.5 * av * av * c2, -0.5 * aw * aw * c2
        L12, L13, L23 = -0.5 * au * av * s2, -0.5 * au * aw * c2, -0.5 * av * aw * s2
        M11, M22, M12, M13 = (-6.0 * d2 * sigma * q for q in (AS * c, -AS * c, AS * s_, S3))
        LM[kk - 1] = L11 * M11 + L22 * M22 + 2.0 * (L12 * M12 + L13 * M13) + 0.0 * L33 + 0.0 * L23
        MM[kk - 1] = 18.0 * d2 * d2 * sigma ** 4
        N13 = 2.0 * d2 * S3 * (sigma - 32.0 * abs(S3))
        Q13 = 0.5 * au * aw
        QN[kk - 1] = 2.0 * Q13 * N13
        NN[kk - 1] = 2.0 * N13 * N13
    return LM, MM, QN, NN


def _close(got, want, rtol=1e-9):
    scale = np.max(np.abs(want))
    np.testing.assert_allclose(got, want, rtol=rtol, atol=rtol * scale)


def _fields(s, p):
    K = np.s_[1:p.nz + 1, :, :p.nx]
    return {n: s.f[n][K] for n in ("F_LM", "F_MM", "F_QN", "F_NN", "Cs_opt2", "Beta",
                                   "Tn_all")}


def test_lagrange_instantaneous_contractions_and_clips():
    """epsi = 1: F_* are the instantaneous contractions (clipped at 1e-24),
    Cs_opt2 follows from them through beta = cs4d / cs2d (exponent
    ln 2 / (ln 4 - ln 2) = 1), beta = 1 at the lid, the clip at 1/8."""
    p, sp, s, t = setup()
    for n in ("F_LM", "F_MM", "F_QN", "F_NN"):
        s.f[n][:] = 1e100
    o.lagrange_Sdep(sp, s)
    g = _fields(s, p)
    LM, MM, QN, NN = expected(p, t)
    zero = 1e-24
    _close(g["F_LM"], np.maximum(zero, LM)[:, None, :] + 0 * g["F_LM"])
    _close(g["F_MM"], MM[:, None, :] + 0 * g["F_MM"])
    _close(g["F_QN"], np.maximum(zero, QN)[:, None, :] + 0 * g["F_QN"])
    _close(g["F_NN"], NN[:, None, :] + 0 * g["F_NN"])
    # Tn = 1.5 Delta 1e-25 is below the 1e-24 floor (lagrange_Sdep.f90:305)
    assert 1.5 * p.delta * 1e-25 < zero
    assert (g["Tn_all"] == zero).all()
    cs2 = np.maximum(zero, np.maximum(zero, LM) / MM)
    cs4 = np.maximum(zero, np.maximum(zero, QN) / NN)
    beta = cs4 / cs2
    beta[-1] = 1.0
    cs = cs2 / np.maximum(beta, 1.0 / 8.0)
    # the branches are all exercised: LM < 0 somewhere (F_LM clipped), beta
    # below and above the 1/8 clip on interior levels
    inner = np.s_[1:p.nz - 1]
    assert (LM[inner] < 0).any() and (LM[inner] > 0).any()
    ok = (cs2 > 1e-6 * cs2.max())
    assert (beta[inner][ok[inner]] < 0.125).any() and (beta[inner][ok[inner]] > 0.125).any()
    _close(g["Cs_opt2"], cs[:, None, :] + 0 * g["Cs_opt2"])
    bg = g["Beta"][:, 0, :]
    np.testing.assert_allclose(bg[ok], beta[ok], rtol=1e-8)
    assert (g["Beta"][-1] == 1.0).all()


def test_lagrange_relaxation_weight_half():
    """F_* = 16 (F_LM F_MM = 256) and lagran_dt = 0.75 Delta: Tn =
    1.5 Delta 256^(-1/8) = 0.75 Delta = lagran_dt, so epsi = 1/2 exactly and
    F_new = (instantaneous + 16) / 2 (lagrange_Sdep.f90:302-313, 345-356)."""
    p, sp, s, t = setup()
    for n in ("F_LM", "F_MM", "F_QN", "F_NN"):
        s.f[n][:] = 16.0
    s.lagran_dt = 0.75 * p.delta
    o.lagrange_Sdep(sp, s)
    g = _fields(s, p)
    LM, MM, QN, NN = expected(p, t)
    zero = 1e-24
    _close(g["F_LM"], np.maximum(zero, 0.5 * LM + 8.0)[:, None, :] + 0 * g["F_LM"])
    _close(g["F_MM"], (0.5 * MM + 8.0)[:, None, :] + 0 * g["F_MM"])
    _close(g["F_QN"], np.maximum(zero, 0.5 * QN + 8.0)[:, None, :] + 0 * g["F_QN"])
    _close(g["F_NN"], (0.5 * NN + 8.0)[:, None, :] + 0 * g["F_NN"])
    np.testing.assert_allclose(g["Tn_all"], 0.75 * p.delta, rtol=1e-14)


def test_lagrange_first_call_initialisation():
    """Fresh start, first dynamic call (jt = cs_count), lagran_dt = 0:
    F_MM = MM, F_LM = 0.03 MM, F_NN = NN, F_QN = 0.03 NN, the padding columns
    1 (lagrange_Sdep.f90:266-277, 326-337), and with epsi = 0 they stay, so
    Cs_opt2 = 0.03 and beta = 1."""
    p, sp, s, t = setup(initu=False)
    s.lagran_dt = 0.0
    o.lagrange_Sdep(sp, s)
    g = _fields(s, p)
    LM, MM, QN, NN = expected(p, t)
    _close(g["F_MM"], MM[:, None, :] + 0 * g["F_MM"])
    _close(g["F_LM"], 0.03 * MM[:, None, :] + 0 * g["F_LM"])
    _close(g["F_NN"], NN[:, None, :] + 0 * g["F_NN"])
    _close(g["F_QN"], 0.03 * NN[:, None, :] + 0 * g["F_QN"])
    np.testing.assert_allclose(g["Cs_opt2"], 0.03, rtol=1e-12)
    np.testing.assert_allclose(g["Beta"], 1.0, rtol=1e-12)
    for n in ("F_LM", "F_MM", "F_QN", "F_NN"):
        assert (s.f[n][1:p.nz + 1, :, p.ld - 2:] == 1.0).all()
    assert s.lm_init and s.qn_init


@pytest.mark.parametrize("wrong", ["const", "tf", "w_bar"])
def test_lagrange_pin_catches_planted_mistakes(wrong):
    """The closed forms above fail on plausible slips planted in a copy of
    the oracle's inputs or constants: const = Delta^2 instead of 2 Delta^2,
    the two filter widths swapped, w_bar taken from w(k+1)."""
    p, sp, s, t = setup()
    for n in ("F_LM", "F_MM", "F_QN", "F_NN"):
        s.f[n][:] = 1e100
    if wrong == "const":
        p.delta = p.delta / math.sqrt(2.0)        # const = 2 delta^2 -> delta^2
    elif wrong == "tf":
        sp.G_test, sp.G_test_test = sp.G_test_test, sp.G_test
    else:
        s.w[2:p.nz] = s.w[3:p.nz + 1].copy()
    o.lagrange_Sdep(sp, s)
    p0, _, _, _ = setup()
    LM, MM, QN, NN = expected(p0, t)
    got = s.f["F_MM"][1:p.nz + 1, 0, :p.nx] if wrong != "w_bar" else \
        s.f["F_QN"][1:p.nz + 1, 0, :p.nx]
    want = MM if wrong != "w_bar" else np.maximum(1e-24, QN)
    assert not np.allclose(got, want, rtol=1e-6, atol=1e-6 * np.max(np.abs(want)))


# --------------------------------------------------------------------------
# trilinear_interp_w: wall cells, lid cell and the periodic seam
# --------------------------------------------------------------------------
def _wgrid_field(p, fz, fx=None, fy=None):
    """var on the interpolation's grid: level 1 on the uv node z = dz/2
    (the wall-model case), levels 2..nz on w nodes z = (k-1) dz."""
    zs = np.array([np.nan, 0.5 * p.dz] + [(k - 1.0) * p.dz for k in range(2, p.nz + 1)])
    x = np.arange(p.nx) * p.dx
    y = np.arange(p.ny) * p.dy
    var = np.zeros((p.nz + 1, p.ny, p.nx))
    for k in range(1, p.nz + 1):
        var[k] = fz(zs[k]) + (fx(x)[None, :] if fx else 0.0) + (fy(y)[:, None] if fy else 0.0)
    return var


def test_interp_near_wall_cells_linear_in_z():
    """Below z(1) = dz/2 the value is var(1) (zero gradient); between z(1)
    and zw(2) = dz it is linear between var(1) at dz/2 and var(2) at dz
    (functions.f90:413-430); at the lid point zw(nz) it is var(nz)."""
    p = o.Params(16, 12, 10)
    f = lambda z: 1.5 - 3.0 * z
    var = _wgrid_field(p, f)
    rng = np.random.default_rng(3)
    n = 64
    px = rng.uniform(0.0, p.lx, n)
    py = rng.uniform(0.0, p.ly, n)
    pz_low = rng.uniform(0.0, 0.5 * p.dz, n)
    pz_mid = rng.uniform(0.5 * p.dz, p.dz, n)
    np.testing.assert_allclose(o.trilinear_interp_w(p, var, px, py, pz_low), f(0.5 * p.dz),
                               rtol=1e-13)
    np.testing.assert_allclose(o.trilinear_interp_w(p, var, px, py, pz_mid), f(pz_mid),
                               rtol=1e-12)
    zlid = np.full(n, (p.nz - 1.0) * p.dz)
    np.testing.assert_allclose(o.trilinear_interp_w(p, var, px, py, zlid), f(zlid), rtol=1e-12)
    pz_top = rng.uniform((p.nz - 2.0) * p.dz, (p.nz - 1.0) * p.dz, n)
    np.testing.assert_allclose(o.trilinear_interp_w(p, var, px, py, pz_top), f(pz_top),
                               rtol=1e-12)


def test_interp_periodic_seam_matches_numpy_interp():
    """Across the last cell in x and y (and from points outside [0, L),
    wrapped with MODULO) the interpolant is numpy's periodic piecewise-linear
    interpolation of the nodal values."""
    p = o.Params(16, 12, 10)
    rng = np.random.default_rng(4)
    gx = rng.standard_normal(p.nx)
    gy = rng.standard_normal(p.ny)
    xs = np.arange(p.nx) * p.dx
    ys = np.arange(p.ny) * p.dy
    var = _wgrid_field(p, lambda z: 0.0 * z, lambda x: gx, lambda y: gy)
    n = 200
    px = np.concatenate([rng.uniform((p.nx - 1) * p.dx, p.lx, n // 2),
                         rng.uniform(-p.lx, 2 * p.lx, n // 2)])
    py = np.concatenate([rng.uniform((p.ny - 1) * p.dy, p.ly, n // 2),
                         rng.uniform(-p.ly, 2 * p.ly, n // 2)])
    pz = rng.uniform(2.0 * p.dz, 5.0 * p.dz, n)
    want = np.interp(px, xs, gx, period=p.lx) + np.interp(py, ys, gy, period=p.ly)
    np.testing.assert_allclose(o.trilinear_interp_w(p, var, px, py, pz), want, rtol=1e-11,
                               atol=1e-12)
