"""The spectral-resident step against the all-physical one
(LESGO_AMD_PHYSICAL_STEP=1): the same flow, and after materialisation the
same physical fields the reference holds after a step (u, v, w with the
disks' high-wavenumber forcing, RHS, RHS_f, p, dpdx, dpdy, dpdz, divt, and
S_ij and Nu_t, which a step without a dynamic-model update leaves to be
made on demand), on a
fresh start with disks through the Euler step, the Smagorinsky phase and
dynamic steps; and members leaving and joining a batch keep their spectra."""
import os

import numpy as np
import pytest

from oracle import lesgo_oracle as o
from gpu_util import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "RHSx", "RHSy", "RHSz", "RHSx_f", "RHSy_f", "RHSz_f", "dpdx",
          "dpdy", "dpdz", "divtx", "divty", "divtz", "fxa", "Cs_opt2", "F_LM", "S11", "S13",
          "S23", "S33", "Nu_t", "txx", "txz")


def _case(physical, precision="fp64", members=1):
    import lesgo_amd
    if physical:
        os.environ["LESGO_AMD_PHYSICAL_STEP"] = "1"
    try:
        cfg = lesgo_amd.Config(48, 32, 16, initu=False, dyn_init=6, precision=precision)
        return lesgo_amd.Case(cfg, members=members)
    finally:
        os.environ.pop("LESGO_AMD_PHYSICAL_STEP", None)


def _levels(name, nz):
    if name in ("u", "v", "RHSx", "RHSy", "RHSx_f", "RHSy_f", "dpdx", "dpdy", "divtx",
                "divty"):
        return np.s_[1:nz]
    if name == "p":
        return np.s_[0:nz + 1]
    return np.s_[1:nz + 1]


@pytest.mark.parametrize("graph", [False, True])
def test_spectral_step_equals_physical_step(graph):
    p = o.Params(48, 32, 16)
    specs = o.aligned_farm(p, 2, 2, dia=0.3, hub=0.3, sx_d=7, sy_d=5)
    rows = [(t.xloc, t.yloc, t.height, t.dia, t.thk, t.theta1, t.theta2, t.ct_prime)
            for t in specs]
    cases = [_case(True), _case(False)]
    try:
        for c in cases:
            c.seed_state(5)
            c.turbines_init(rows)
        for chunk in (1, 4, 3, 5):
            for c in cases:
                (c.step_graph if graph else c.step)(chunk)
            worst = 0.0
            for n in FIELDS:
                K = _levels(n, 17)
                a = cases[0].download(n)[K][:, :, :48]
                b = cases[1].download(n)[K][:, :, :48]
                e = rel_l2(b, a)
                worst = max(worst, e)
                assert e <= 1e-12, (n, e)
            print(f"  step {cases[0].counters()[1]}: worst {worst:.2e}", flush=True)
    finally:
        for c in cases:
            c.close()


def test_spectral_members_leave_and_join():
    """A 3-member batch where member 1 is retired mid-run and member 0 is
    re-seeded: every member's fields equal the same member stepped alone,
    bit for bit (the members' spectra move with them)."""
    import lesgo_amd
    cfg = lesgo_amd.Config(32, 32, 16, initu=False, dyn_init=6)
    alone = {}
    for seed, steps in ((1, 9), (2, 12), (3, 4)):
        with lesgo_amd.Case(cfg) as a:
            a.seed_state(seed)
            a.step_graph(steps)
            alone[seed] = {n: a.download(n) for n in ("u", "v", "w", "p", "RHSx_f")}
    with lesgo_amd.Case(cfg, members=3) as b:
        for m, seed in enumerate((1, 2, 7)):
            b.seed_state(seed, member=m)
        b.step_graph(5)
        b.seed_state(3, member=2)        # member 2 restarts (alone: 4 steps)
        b.step_graph(4)
        got1 = {n: b.download(n, 0) for n in ("u", "v", "w", "p", "RHSx_f")}
        got3 = {n: b.download(n, 2) for n in ("u", "v", "w", "p", "RHSx_f")}
        b.copy_member(2, 0)              # member 0 := member 2
        b.step_graph(3)
        got2 = {n: b.download(n, 1) for n in ("u", "v", "w", "p", "RHSx_f")}
    for got, seed in ((got1, 1), (got2, 2), (got3, 3)):
        for n, r in alone[seed].items():
            assert np.array_equal(got[n][1:17], r[1:17]), (seed, n)
