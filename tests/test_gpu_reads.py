"""What a case's fields read back as must not depend on when the host looked.

The spectral-resident step makes the physical fields on demand; the part of
RHS_f its spectra do not carry (the high-wavenumber content of the previous
step's disk forcing and vertical stresses) comes from the previous step's
stresses and forcing, or, for a member whose step started from its physical
fields, from its physical RHS_f.  These tests hold:

* a run with a download after every step equals the same run without reads,
  bit for bit, with tilted and yawed disks (their uv-grid vertical forcing
  has its own buffer, untouched by the reads' scratch use);
* a disk set replaced mid-run (other positions, tilt, then none) matches
  the oracle at 1e-10 with and without reads (the forcing buffers are
  cleared as they are rewritten, the previous step's forcing kept for the
  physical fields);
* a running member of a batch is unchanged, bit for bit, by another member
  joining (seeded, with disks) for its last step.
"""
import numpy as np
import pytest

from oracle import lesgo_oracle as o
from gpu_util import gpu_case, rel_l2, upload_all

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "RHSx_f", "RHSz_f", "fza")


def tilted_farm(dia=0.3, hub=0.3):
    # (xloc, yloc, height, dia, thk, theta1 yaw, theta2 tilt, Ct')
    return [(1.6, 0.8, hub, dia, 0.0, 20.0, 10.0, 1.33),
            (4.2, 2.3, hub, dia, 0.0, -15.0, 5.0, 1.2)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reads_do_not_change_results(precision):
    import lesgo_amd
    out = []
    for reads in (False, True):
        cfg = lesgo_amd.Config(48, 32, 24, precision=precision, initu=False, dyn_init=5)
        with lesgo_amd.Case(cfg) as c:
            c.seed_state(3)
            c.turbines_init(tilted_farm())
            for _ in range(8):
                c.step_graph(1)
                if reads:
                    c.download("u")
                    c.download("fza")
            out.append({n: c.download(n) for n in FIELDS})
    for n in FIELDS:
        d = float(np.max(np.abs(out[0][n][1:24] - out[1][n][1:24])))
        print(f"  {precision} {n}: max |reads - none| = {d:.1e}", flush=True)
        assert d == 0.0, (n, d)


def _specs(lst):
    return [o.TurbineSpec(*t) for t in lst]


@pytest.mark.parametrize("reads", [False, True])
def test_disk_set_changes_match_oracle(reads):
    p = o.Params(48, 32, 24)
    sp = o.Spectral(p)
    s = o.initial_condition(p, 0)
    farm_a = [(t.xloc, t.yloc, t.height, t.dia, t.thk, t.theta1, t.theta2, t.ct_prime)
              for t in o.aligned_farm(p, 2, 2, dia=0.3, hub=0.3, sx_d=7, sy_d=5)]
    farm_b = tilted_farm()
    plan = {0: farm_a, 3: farm_b, 6: []}
    with gpu_case(p) as c:
        upload_all(c, s)
        for k in range(9):
            if k in plan:
                s.turbines = o.Turbines(p, _specs(plan[k])) if plan[k] else None
                c.turbines_init(plan[k])
            o.step(sp, s)
            c.step_graph(1)
            if reads or k in (2, 5, 8):
                X = np.s_[:, :, :p.nx]
                e = {n: rel_l2(c.download(n)[lv][X], s.f[n][lv][X])
                     for n, lv in (("u", np.s_[1:p.nz]), ("v", np.s_[1:p.nz]),
                                   ("w", np.s_[1:p.nz + 1]), ("p", np.s_[0:p.nz + 1]),
                                   ("RHSx_f", np.s_[1:p.nz]), ("RHSz_f", np.s_[1:p.nz]))}
                print(f"  step {k + 1}: " + " ".join(f"{a}={b:.1e}" for a, b in e.items()),
                      flush=True)
                for n, v in e.items():
                    assert v <= 1e-10, (k + 1, n, v)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_member_joining_leaves_running_member_unchanged(precision):
    import lesgo_amd
    cfg = lesgo_amd.Config(96, 48, 64, zo=5e-4, precision=precision, initu=False, dyn_init=10)
    n = 16
    with lesgo_amd.Case(cfg) as a:
        a.seed_state(13)
        a.step_graph(n)
        ref = {f: a.download(f) for f in ("u", "v", "w", "p")}
    farm = lesgo_amd.aligned_farm(cfg, 2, 2, dia=0.15, hub=0.1, sx_d=7.0, sy_d=5.0,
                                  ct_prime=1.33)
    with lesgo_amd.Case(cfg, members=2) as b:
        b.set_active(1)
        b.seed_state(13, member=0)
        b.step_graph(n - 1)
        b.set_member(1, 1e-5, 1.0, 0.0)
        b.seed_state(18, member=1)
        b.turbines_init(farm, member=1)
        b.set_active(2)
        b.step_graph(1)
        for f, r in ref.items():
            d = float(np.max(np.abs(b.download(f, 0)[1:65] - r[1:65])))
            print(f"  {precision} {f}: max |joined - alone| = {d:.1e}", flush=True)
            assert d == 0.0, (f, d)
