"""Pins the oracle's running time averages (time_average.f90 tavg%compute,
io.f90:558-582 driving) and the project-defined streamwise spectra."""
import math

import numpy as np

from oracle import lesgo_oracle as o


def setup(nx=16, ny=12, Nz=8):
    p = o.Params(nx, ny, Nz)
    return p, o.Spectral(p), o.State.zeros(p)


def drive(st, sp, s, jts):
    for jt in jts:
        s.jt_total = jt
        st.after_step(sp, s)


def test_window_and_weights():
    """nstart 2, nskip 2, nend 9: the call at 2 starts the sums, 4, 6, 8 add
    2 dt each; 9 adds dt to the pending weight only."""
    p, sp, s = setup()
    st = o.Stats(p, 2, 9, 2)
    s.u[:] = 3.0
    drive(st, sp, s, range(1, 5))
    s.u[:] = 5.0
    drive(st, sp, s, range(5, 12))
    assert math.isclose(st.total_time, 6 * p.dt, rel_tol=1e-15)
    # u = 3 at jt 4 (weight 2dt), 5 at jt 6, 8 (weight 4dt)
    np.testing.assert_allclose(st.mean("u"), (3.0 * 2 + 5.0 * 4) / 6, rtol=1e-14)
    assert math.isclose(st.dt, p.dt, rel_tol=1e-15)


def test_interpolations_exact_on_linear_profiles():
    p, sp, s = setup()
    st = o.Stats(p, 1, 3, 1)
    zu = (np.arange(p.nz + 1) - 0.5) * p.dz           # uv nodes
    zw = (np.arange(p.nz + 1) - 1.0) * p.dz           # w nodes
    s.u[:] = (2.0 * zu + 1.0)[:, None, None]
    s.v[:] = (-1.5 * zu)[:, None, None]
    s.w[:] = (0.7 * zw + 0.2)[:, None, None]
    drive(st, sp, s, (1, 2))
    K = np.s_[2:p.nz + 1]
    np.testing.assert_allclose(st.mean("u_w")[K][:, 0, 0], 2.0 * zw[K] + 1.0, rtol=1e-13)
    np.testing.assert_allclose(st.mean("v_w")[K][:, 0, 0], -1.5 * zw[K], rtol=1e-13)
    assert (st.mean("u_w")[1] == 0).all() and (st.mean("v_w")[1] == 0).all()
    J = np.s_[1:p.nz]
    np.testing.assert_allclose(st.mean("w_uv")[J][:, 0, 0], 0.7 * zu[J] + 0.2, rtol=1e-13)
    np.testing.assert_allclose(st.mean("w_uv")[p.nz], st.mean("w_uv")[p.nz - 1])
    np.testing.assert_allclose(st.mean("uw")[K][:, 0, 0],
                               (2.0 * zw[K] + 1.0) * (0.7 * zw[K] + 0.2), rtol=1e-13)


def test_pressure_and_vorticity_closed_forms():
    p, sp, s = setup()
    st = o.Stats(p, 1, 2, 1)
    s.u[:] = 4.0
    s.v[:] = -1.0
    s.f["p"][:] = 10.0
    s.dudz[:] = 2.5
    s.dwdx[:] = 0.5
    s.dvdx[:] = 1.25
    drive(st, sp, s, (1, 2))
    J = np.s_[1:p.nz]
    np.testing.assert_allclose(st.mean("p")[J], 10.0 - 0.5 * (16.0 + 1.0), rtol=1e-14)
    np.testing.assert_allclose(st.mean("vorty"), 2.0, rtol=1e-14)
    np.testing.assert_allclose(st.mean("vortz")[2:], 1.25, rtol=1e-14)
    assert (st.mean("vortz")[:2] == 0).all()


def test_spectrum_single_mode_and_parseval():
    p, sp, s = setup(nx=32, ny=16, Nz=6)
    x = np.arange(p.nx) * p.lx / p.nx
    A, B, m = 1.7, 0.4, 3
    s.u[:, :, :p.nx] = B + A * np.cos(2 * np.pi * m * x / p.lx)
    rng = np.random.default_rng(1)
    s.v[:] = rng.standard_normal(s.v.shape)
    s.v[:], _, _ = o.filt_da(sp, s.v)           # no Nyquist modes left
    st = o.Stats(p, 1, 2, 1)
    drive(st, sp, s, (1, 2))
    E = st.spectrum(0)
    K = np.s_[1:p.nz]
    np.testing.assert_allclose(E[K][:, 0], B * B, rtol=1e-13)
    np.testing.assert_allclose(E[K][:, m], A * A / 2, rtol=1e-13)
    others = np.delete(E[K], [0, m], axis=1)
    assert np.abs(others).max() < 1e-25
    Ev = st.spectrum(1)
    np.testing.assert_allclose(Ev[K].sum(axis=1), (s.v[K][:, :, :p.nx] ** 2).mean(axis=(1, 2)),
                               rtol=1e-12)
