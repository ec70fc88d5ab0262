"""Pins for the oracle's CFL-adaptive time step (cfl_util.f90:73-111,
main.f90:135-144, initialize.f90:191-200, sgs_stag_util.f90:74-82,
lagrange_Sdep.f90:430), each against something the formula does not restate:
a hand-placed velocity maximum, the exactness of variable-step AB2 for a
right-hand side linear in time, the Euler limit, and lagran_dt as a
difference of total times."""
import math

import numpy as np
import pytest

from oracle import lesgo_oracle as o


def _state(p):
    s = o.State.zeros(p)
    for n in ("u", "v", "w"):
        s.f[n][1:p.nz, :, :p.nx] = 0.1
    return s


@pytest.mark.parametrize("which", ["u", "v", "w"])
def test_get_cfl_dt_hand_placed_maximum(which):
    """dt = cfl / (|peak| / h) for a single peak on levels 1..nz-1, i < nx;
    larger values on level 0, level nz and the padding columns are outside
    the range and must not count."""
    p = o.Params(16, 12, 8, use_cfl_dt=True, cfl=0.3)
    s = _state(p)
    h = {"u": p.dx, "v": p.dy, "w": p.dz}[which]
    peak = -7.25
    s.f[which][3, 5, 9] = peak
    for n in ("u", "v", "w"):
        s.f[n][0] = 1e3
        s.f[n][p.nz] = 1e3
        s.f[n][1:p.nz, :, p.nx:] = 1e3
    assert o.get_cfl_dt(p, s) == pytest.approx(p.cfl * h / abs(peak), rel=1e-15)
    # the same peak on level nz (excluded) leaves the background maximum
    s2 = _state(p)
    s2.f[which][p.nz, 5, 9] = peak
    hmin = min(p.dx, p.dy, p.dz)
    assert o.get_cfl_dt(p, s2) == pytest.approx(p.cfl * hmin / 0.1, rel=1e-15)


def test_variable_step_ab2_exact_for_linear_rhs():
    """Driving cfl_clock with a velocity that sets a different dt each step,
    u_{n+1} = u_n + dt (tadv1 f_n + tadv2 f_{n-1}) integrates f(t) = a + b t
    exactly from the second step on (variable-step AB2 is exact on linear f);
    the Euler first step (tadv1 = 1, tadv2 = 0) misses b dt_1^2 / 2."""
    p = o.Params(8, 8, 4, use_cfl_dt=True, cfl=0.5)
    s = _state(p)
    a, b = 0.7, -1.3
    f = lambda t: a + b * t
    t, y, f_prev = 0.0, 0.0, None
    rng = np.random.default_rng(3)
    for n in range(12):
        umax = float(rng.uniform(0.5, 3.0))
        s.f["u"][1:p.nz, :, :p.nx] = 0.0
        s.f["u"][1, 0, 0] = umax
        s.f["v"][:] = 0.0
        s.f["w"][:] = 0.0
        o.cfl_clock(p, s)
        assert s.dt == pytest.approx(p.cfl * p.dx / umax, rel=1e-15)
        if n == 0:
            assert s.tadv1 == 1.0 and s.tadv2 == 0.0
            f_prev = f(t)
            dt1 = s.dt
        y = y + s.dt * (s.tadv1 * f(t) + s.tadv2 * f_prev)
        f_prev = f(t)
        t = t + s.dt
        exact = a * t + 0.5 * b * t * t - 0.5 * b * dt1 * dt1
        assert y == pytest.approx(exact, rel=1e-12, abs=1e-15), n


def test_first_cfl_step_is_the_euler_step_of_that_dt():
    """A fresh CFL clock's first step equals a fixed-dt fresh start run with
    dt = get_cfl_dt of the initial field and the Euler weights tadv1 = 1,
    tadv2 = 0 (the pressure solve and the projection see tadv1: the Neumann
    rows from divtz are not scaled by it, press_stag_array.f90:52,114-124)."""
    pc = o.Params(16, 16, 8, use_cfl_dt=True, cfl=0.05, initu=False, dyn_init=100)
    s = o.initial_condition(pc, 4)
    dt0 = o.get_cfl_dt(pc, s)
    pf = o.Params(16, 16, 8, dt=dt0, initu=False, dyn_init=100)
    pf.tadv1, pf.tadv2 = 1.0, 0.0
    sf = o.initial_condition(pf, 4)
    o.step(o.Spectral(pc), s)
    o.step(o.Spectral(pf), sf)
    assert s.dt == dt0 and s.total_time == dt0
    for n in ("u", "v", "w", "p"):
        ref = sf.f[n][1:pf.nz, :, :pf.nx]
        err = np.linalg.norm(s.f[n][1:pc.nz, :, :pc.nx] - ref) / np.linalg.norm(ref)
        assert err < 1e-13, (n, err)


def test_lagran_dt_is_time_since_last_dynamic_call(monkeypatch):
    """lagran_dt seen by the dynamic model equals total_time now minus
    total_time at the previous dynamic call (file start), and for a fresh
    start the time since step DYN_init - cs_count."""
    seen = []
    orig = o.interpolag_Sdep

    def spy(p, s):
        seen.append((s.jt_total, s.lagran_dt, s.total_time))
        return orig(p, s)

    monkeypatch.setattr(o, "interpolag_Sdep", spy)
    for initu, dyn_init in ((True, 100), (False, 10)):
        seen.clear()
        p = o.Params(16, 16, 8, use_cfl_dt=True, cfl=0.05, initu=initu, dyn_init=dyn_init)
        sp = o.Spectral(p)
        s = o.initial_condition(p, 1)
        times = {0: 0.0}
        for _ in range(15):
            o.step(sp, s)
            times[s.jt_total] = s.total_time
        calls = [c[0] for c in seen]
        assert calls == ([5, 10, 15] if initu else [10, 15]), calls
        for jt, ldt, tt in seen:
            start = jt - p.cs_count
            assert ldt == pytest.approx(tt - times[start], rel=1e-12), (initu, jt)
        assert s.lagran_dt == 0.0
