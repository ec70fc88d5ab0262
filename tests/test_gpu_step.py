"""Whole-step parity: 100 steps of the 32^3 neutral half channel (dt = 2e-4,
seed 0), compared with the oracle every 10 steps.  Bar: 1e-10 relative L2 on
u, v, w and p over their valid ranges."""
import numpy as np
import pytest

from oracle import lesgo_oracle as o
from gpu_util import gpu_case, rel_l2

pytestmark = pytest.mark.gpu


def run_pair(p, nsteps, every, seed=0, turbines=None, precision="fp64"):
    import lesgo_amd
    sp = o.Spectral(p)
    s = o.initial_condition(p, seed)
    if turbines is not None:
        s.turbines = o.Turbines(p, turbines)
    errs = []
    with gpu_case(p, precision) as c:
        for n, a in s.f.items():
            c.upload(n, a)
        c.set_counters(0, 0)
        if turbines is not None:
            c.turbines_init([(t.xloc, t.yloc, t.height, t.dia, t.thk, t.theta1, t.theta2, t.ct_prime)
                             for t in turbines])
        for k in range(nsteps // every):
            for _ in range(every):
                o.step(sp, s)
            c.step(every)
            e = {}
            X = np.s_[:, :, :p.nx]
            e["u"] = rel_l2(c.download("u")[1:p.nz][X], s.u[1:p.nz][X])
            e["v"] = rel_l2(c.download("v")[1:p.nz][X], s.v[1:p.nz][X])
            e["w"] = rel_l2(c.download("w")[1:p.nz + 1][X], s.w[1:p.nz + 1][X])
            e["p"] = rel_l2(c.download("p")[0:p.nz + 1][X], s.f["p"][0:p.nz + 1][X])
            e["rmsdiv_gpu"] = c.rmsdiv()
            e["rmsdiv_ref"] = o.rmsdiv(p, s)
            print(f"  step {s.jt_total}: " + " ".join(f"{kk}={vv:.2e}" for kk, vv in e.items()),
                  flush=True)
            errs.append(e)
        extra = c.turbines_state() if turbines is not None else None
    return errs, s, extra


def test_config0_100_steps_fp64():
    p = o.Params(32, 32, 32)
    errs, s, _ = run_pair(p, 100, 10)
    worst = max(max(e[k] for k in ("u", "v", "w", "p")) for e in errs)
    assert worst <= 1e-10, worst
    # rmsdiv every 10 steps (wbase = 10): roundoff of the projection in both
    for e in errs:
        assert abs(e["rmsdiv_gpu"] - e["rmsdiv_ref"]) <= 1e-10 * e["rmsdiv_ref"] + 1e-12, e
    assert errs[-1]["rmsdiv_gpu"] < 1e-12


def test_fresh_start_fp64():
    """Fresh start (initu = .false.): Euler first step, Smagorinsky phase,
    F_* initialisation at DYN_init = 10 and dynamic steps 15, 20: u, v, w, p
    and Cs_opt2, F_LM against the oracle at 1e-10."""
    import lesgo_amd
    p = o.Params(32, 32, 16, initu=False, dyn_init=10)
    sp = o.Spectral(p)
    s = o.initial_condition(p, 2)
    with gpu_case(p) as c:
        for n, a in s.f.items():
            c.upload(n, a)
        c.set_counters(0, 0)
        for chunk in (1, 8, 1, 5, 6):
            for _ in range(chunk):
                o.step(sp, s)
            c.step(chunk)
            X = np.s_[1:p.nz, :, :p.nx]
            errs = {n: rel_l2(c.download(n)[X], s.f[n][X])
                    for n in ("u", "v", "w", "p", "Cs_opt2", "F_LM", "F_NN")}
            print(f"  step {s.jt_total}: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()),
                  flush=True)
            for n, e in errs.items():
                assert e <= 1e-10, (s.jt_total, n, e)
    assert lesgo_amd


def test_config0_mixed_precision_error():
    """Mixed step (FP32 FFT/convection/SGS, FP64 tridiagonal) against the FP64
    oracle: velocity rel L2 <= 1e-4 after 100 steps."""
    p = o.Params(32, 32, 32)
    errs, s, _ = run_pair(p, 100, 50, precision="mixed")
    for k in ("u", "v", "w"):
        assert errs[-1][k] <= 1e-4, (k, errs[-1][k])


def test_turbines_step_fp64():
    """Aligned 2x2 actuator-disk array on a 48x32x24 grid: u_d_T, thrust and
    the flow after 9 steps (forcing every step, one dynamic update) match the
    oracle at 1e-10."""
    p = o.Params(48, 32, 24)
    specs = o.aligned_farm(p, 2, 2, dia=0.3, hub=0.3, sx_d=7, sy_d=5)
    errs, s, (udt, fn) = run_pair(p, 9, 3, turbines=specs)
    np.testing.assert_allclose(udt, s.turbines.u_d_T, rtol=1e-10)
    np.testing.assert_allclose(fn, s.turbines.f_n, rtol=1e-10)
    worst = max(max(e[k] for k in ("u", "v", "w", "p")) for e in errs)
    assert worst <= 1e-10, worst


def test_turbines_long_horizon_bounded():
    """Past the step-10 dynamic update the reference's Lagrangian timescale
    Tn ~ max(F_LM F_MM, 1e-24)^(-1/8) (lagrange_Sdep.f90:301-303) is
    hypersensitive where F_LM sits at its clip floor near the lid: the oracle
    fed 1e-15 relative noise per step departs from itself by ~1e-7..1e-6 by
    step 20 (tests/test_oracle_ops.py documents the effect).  Beyond that
    point the bar is a bug-catching absolute bound of 1e-4 relative L2; the
    oracle's own noise response is printed beside it."""
    p = o.Params(48, 32, 24)
    specs = o.aligned_farm(p, 2, 2, dia=0.3, hub=0.3, sx_d=7, sy_d=5)
    errs, s, _ = run_pair(p, 20, 10, turbines=specs)
    sp = o.Spectral(p)
    a = o.initial_condition(p, 0)
    a.turbines = o.Turbines(p, specs)
    b = a.copy()
    rng = np.random.default_rng(0)
    for _ in range(20):
        o.step(sp, a)
        o.step(sp, b)
        for f in ("u", "v", "w"):
            b.f[f][1:p.nz] *= 1 + 1e-15 * rng.standard_normal(b.f[f][1:p.nz].shape)
    X = np.s_[1:p.nz, :, :p.nx]
    for k in ("u", "v", "w"):
        self_dev = rel_l2(b.f[k][X], a.f[k][X])
        print(f"  {k}: gpu-oracle {errs[-1][k]:.2e}  oracle noise response {self_dev:.2e}")
        assert errs[-1][k] <= 1e-4
