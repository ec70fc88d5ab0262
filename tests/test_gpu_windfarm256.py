"""The benchmark workload itself against the oracle: the wind farm at
256 x 256 x 192 (Lx = 2 pi, Ly = pi, Lz = 1, dt = 1e-4, z0 = 1e-4), a 4 x 4
aligned actuator-disk array (D = 0.1, hub 0.1, 7D x 5D, Ct' = 1.33,
Gaussian-filtered indicator, ADM correction), seeded log-law start (seed 7)
run as a fresh start with DYN_init = 5: the Euler first step, three
Smagorinsky steps and, at step 5, the F_* initialisation and the first
Lagrangian dynamic update, with the disk forcing on every step.

FP64: u, v, w, p, Cs_opt2 and F_LM at 1e-10 relative L2 after steps 1 and 5,
and each disk's u_d_T and thrust at 1e-10.  Mixed precision against the same
FP64 oracle: velocity within 1e-5 (the rounding of float storage and
transforms, not a drift).  The oracle takes about a minute for the five
steps at this size."""
import numpy as np
import pytest

from oracle import lesgo_oracle as o
from gpu_util import gpu_case, rel_l2, upload_all

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

NX, NY, NZ = 256, 256, 192
CHECK_AT = (1, 5)


@pytest.fixture(scope="module")
def reference():
    p = o.Params(NX, NY, NZ, dt=1e-4, initu=False, dyn_init=5)
    sp = o.Spectral(p)
    s = o.initial_condition(p, 7)
    s.turbines = o.Turbines(p, o.aligned_farm(p))
    start = s.copy()
    start.turbines = s.turbines.copy()
    snaps = {}
    for _ in range(max(CHECK_AT)):
        o.step(sp, s)
        print(f"  oracle step {s.jt_total} done", flush=True)
        if s.jt_total in CHECK_AT:
            snaps[s.jt_total] = {
                "f": {n: s.f[n][1:p.nz + 1].copy()
                      for n in ("u", "v", "w", "p", "Cs_opt2", "F_LM")},
                "u_d_T": s.turbines.u_d_T.copy(), "f_n": s.turbines.f_n.copy(),
                "rmsdiv": o.rmsdiv(p, s)}
    return p, start, snaps


def _run(p, start, snaps, precision):
    specs = [(t.xloc, t.yloc, t.height, t.dia, t.thk, t.theta1, t.theta2, t.ct_prime)
             for t in start.turbines.specs]
    out = {}
    with gpu_case(p, precision) as c:
        upload_all(c, start)
        c.turbines_init(specs)
        done = 0
        for k in CHECK_AT:
            c.step(k - done)
            done = k
            ref = snaps[k]
            e = {}
            for n, r in ref["f"].items():
                lv = slice(0, p.nz - 1) if n != "w" else slice(0, p.nz)
                e[n] = rel_l2(c.download(n)[1:p.nz + 1][lv, :, :p.nx], r[lv, :, :p.nx])
            udt, fn = c.turbines_state()
            e["u_d_T"] = float(np.max(np.abs(udt - ref["u_d_T"]) / np.abs(ref["u_d_T"])))
            e["f_n"] = float(np.max(np.abs(fn - ref["f_n"]) / np.abs(ref["f_n"])))
            e["rmsdiv_gpu"] = c.rmsdiv()
            e["rmsdiv_oracle"] = ref["rmsdiv"]
            print(f"  {precision} step {k}: " + " ".join(f"{a}={b:.2e}" for a, b in e.items()),
                  flush=True)
            out[k] = e
    return out


def test_windfarm256_fp64_matches_oracle(reference):
    p, start, snaps = reference
    errs = _run(p, start, snaps, "fp64")
    for k, e in errs.items():
        for n in ("u", "v", "w", "p", "Cs_opt2", "F_LM", "u_d_T", "f_n"):
            assert e[n] <= 1e-10, (k, n, e[n])
        # rmsdiv uses the start-of-step derivatives: the seeded field's
        # divergence at step 1, roundoff after a projection later
        assert abs(e["rmsdiv_gpu"] - e["rmsdiv_oracle"]) <= 1e-10 * e["rmsdiv_oracle"] + 1e-12


def test_windfarm256_mixed_against_fp64_oracle(reference):
    p, start, snaps = reference
    errs = _run(p, start, snaps, "mixed")
    for k, e in errs.items():
        for n in ("u", "v", "w"):
            assert e[n] <= 1e-5, (k, n, e[n])
        assert e["u_d_T"] <= 1e-5
