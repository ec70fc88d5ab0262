"""Batched cases, graph-replayed steps and the autotuner.

A batch member must compute exactly what the same case computes alone
(the kernels, launch choices and arithmetic per member are identical), so
those comparisons are bit-exact.  Graph replay launches the same kernels as
eager stepping: bit-exact too.  A tuned launch choice may change the FFT
split and hence rounding: tuned vs default within 1e-12 (FP64, few steps).
The batched mixed-precision ensemble is checked per member against the FP64
oracle of that member."""
import numpy as np
import pytest

from oracle import lesgo_oracle as o
from gpu_util import rel_l2

pytestmark = pytest.mark.gpu

FIELDS_CMP = ("u", "v", "w", "p", "F_LM", "Cs_opt2", "RHSx")


def cfg_of(nx=32, ny=32, Nz=16, precision="fp64", **kw):
    import lesgo_amd
    return lesgo_amd.Config(nx, ny, Nz, precision=precision, **kw)


MEMBERS = [  # seed, zo, mean_p_force_x, turbines
    (0, 1e-4, 1.0, False),
    (3, 3e-5, 0.7, True),
    (5, 8e-4, 1.3, False),
]


def farm(cfg):
    import lesgo_amd
    return lesgo_amd.aligned_farm(cfg, 2, 2, dia=0.3, hub=0.3, sx_d=7, sy_d=5)


def setup_member(c, m, cfg, seed, zo, mpf, turb):
    c.set_member(m, zo, mpf, 0.0)
    c.seed_state(seed, 0.1, member=m)
    if turb:
        c.turbines_init(farm(cfg), member=m)


def alone(cfg, seed, zo, mpf, turb, nsteps, graph=False):
    import lesgo_amd
    with lesgo_amd.Case(cfg) as c:
        setup_member(c, 0, cfg, seed, zo, mpf, turb)
        c.set_counters(0, 0)
        (c.step_graph if graph else c.step)(nsteps)
        return {n: c.download(n) for n in FIELDS_CMP}, c.rmsdiv()


def test_batch_members_equal_alone_bitwise():
    import lesgo_amd
    cfg = cfg_of()
    nsteps = 7        # crosses the dynamic-model steps 5
    with lesgo_amd.Case(cfg, members=len(MEMBERS)) as c:
        for m, (seed, zo, mpf, turb) in enumerate(MEMBERS):
            setup_member(c, m, cfg, seed, zo, mpf, turb)
        c.set_counters(0, 0)
        c.step(nsteps)
        got = [{n: c.download(n, member=m) for n in FIELDS_CMP} for m in range(len(MEMBERS))]
        rms = [c.rmsdiv(m) for m in range(len(MEMBERS))]
    for m, (seed, zo, mpf, turb) in enumerate(MEMBERS):
        ref, rref = alone(cfg, seed, zo, mpf, turb, nsteps)
        for n in FIELDS_CMP:
            d = np.max(np.abs(got[m][n] - ref[n]))
            print(f"  member {m} {n}: max |batch - alone| = {d:.3e}", flush=True)
            assert d == 0.0, (m, n, d)
        assert rms[m] == rref
    # the members really differ from one another
    assert np.max(np.abs(got[0]["u"] - got[2]["u"])) > 1e-3


def test_batch_member_matches_oracle():
    """Member 1 (own z0, mean pressure gradient, seed and disks) of a batch
    against the FP64 oracle of that case: 1e-10."""
    import lesgo_amd
    seed, zo, mpf, _ = MEMBERS[1]
    p = o.Params(32, 32, 16, zo=zo, mean_p_force_x=mpf)
    sp = o.Spectral(p)
    s = o.initial_condition(p, seed)
    cfg = cfg_of()
    with lesgo_amd.Case(cfg, members=3) as c:
        for m, (sd, z, f, _) in enumerate(MEMBERS):
            c.set_member(m, z, f, 0.0)
            c.seed_state(sd, 0.1, member=m)
        for n, a in s.f.items():
            c.upload(n, a, member=1)
        c.set_counters(0, 0)
        for _ in range(6):
            o.step(sp, s)
        c.step(6)
        X = np.s_[:, :, :p.nx]
        for n, lv in (("u", np.s_[1:p.nz]), ("v", np.s_[1:p.nz]), ("w", np.s_[1:p.nz + 1]),
                      ("p", np.s_[0:p.nz + 1])):
            e = rel_l2(c.download(n, member=1)[lv][X], s.f[n][lv][X])
            print(f"  {n}: rel L2 = {e:.3e}", flush=True)
            assert e <= 1e-10, (n, e)


@pytest.mark.parametrize("precision,fresh", [("fp64", False), ("mixed", False),
                                             ("mixed", True)])
def test_graph_replay_equals_eager(precision, fresh):
    import lesgo_amd
    cfg = cfg_of(precision=precision, initu=not fresh, dyn_init=5)
    nsteps = 12       # all four (parity, dynamic-model) graphs, replayed
    with lesgo_amd.Case(cfg, members=2) as a, lesgo_amd.Case(cfg, members=2) as b:
        for c in (a, b):
            for m in range(2):
                seed, zo, mpf, turb = MEMBERS[m + 1]
                setup_member(c, m, cfg, seed, zo, mpf, turb)
            c.set_counters(0, 0)
        a.step(nsteps)
        for _ in range(nsteps):
            b.step_graph(1)
            b.synchronize()     # eager host work between replays
        for m in range(2):
            for n in FIELDS_CMP:
                d = np.max(np.abs(a.download(n, member=m) - b.download(n, member=m)))
                assert d == 0.0, (precision, m, n, d)
        assert a.counters() == b.counters()


def test_set_active_and_copy_member():
    """Compaction as the serving loop does it: member 2 copied into slot 0,
    one member stepped; it continues exactly as the uninterrupted case."""
    import lesgo_amd
    cfg = cfg_of()
    seed, zo, mpf, turb = MEMBERS[1]
    ref, _ = alone(cfg, seed, zo, mpf, turb, 10)
    with lesgo_amd.Case(cfg, members=3) as c:
        for m, mem in enumerate([MEMBERS[0], MEMBERS[2], MEMBERS[1]]):
            setup_member(c, m, cfg, *mem)
        c.set_counters(0, 0)
        c.step_graph(5)
        c.copy_member(2, 0)
        c.turbines_init([], member=2)
        c.set_active(1)
        assert c.batch_info() == (3, 1)
        c.step_graph(5)
        for n in FIELDS_CMP:
            d = np.max(np.abs(c.download(n, member=0) - ref[n]))
            assert d == 0.0, (n, d)


def test_autotune_records_choices_and_stays_close():
    import lesgo_amd
    from lesgo_amd.case import tune_clear, tune_get
    cfg = cfg_of(48, 32, 16)
    tune_clear()
    ref, _ = alone(cfg, 2, 1e-4, 1.0, False, 6)
    with lesgo_amd.Case(cfg, members=2) as c:
        n = c.autotune(3)
        assert n == 9
    for kind, N, other in (("r2c", 48, 32), ("c2r", 72, 48), ("cols_fwd", 32, 24),
                           ("cols_pad", 48, 24), ("tridag", 32, 24)):
        print(f"  {kind} N={N} other={other}: {tune_get(kind, 'fp64', N, other)}", flush=True)
    got, _ = alone(cfg, 2, 1e-4, 1.0, False, 6)
    tune_clear()
    for n in ("u", "v", "w", "p"):
        e = rel_l2(got[n][1:16, :, :48], ref[n][1:16, :, :48])
        print(f"  tuned vs default {n}: {e:.2e}", flush=True)
        assert e <= 1e-12, (n, e)


def test_mixed_ensemble_members_vs_fp64_oracle():
    """Four-member mixed-precision batch (own seeds, z0, pressure gradients;
    one with disks): each member against the FP64 oracle run of that case
    after 20 steps, velocity rel L2 <= 1e-4."""
    import lesgo_amd
    members = [(0, 1e-4, 1.0), (1, 1e-5, 0.5), (2, 1e-3, 1.5), (3, 3e-4, 0.9)]
    cfg = cfg_of(precision="mixed")
    nsteps = 20
    with lesgo_amd.Case(cfg, members=len(members)) as c:
        refs = []
        for m, (seed, zo, mpf) in enumerate(members):
            p = o.Params(32, 32, 16, zo=zo, mean_p_force_x=mpf)
            s = o.initial_condition(p, seed)
            c.set_member(m, zo, mpf, 0.0)
            for n, a in s.f.items():
                c.upload(n, a, member=m)
            refs.append((p, s))
        c.set_counters(0, 0)
        c.step_graph(nsteps)
        for m, (p, s) in enumerate(refs):
            sp = o.Spectral(p)
            for _ in range(nsteps):
                o.step(sp, s)
            X = np.s_[1:p.nz, :, :p.nx]
            errs = {n: rel_l2(c.download(n, member=m)[X], s.f[n][X]) for n in ("u", "v", "w")}
            print(f"  member {m}: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()),
                  flush=True)
            for n, e in errs.items():
                assert e <= 1e-4, (m, n, e)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_staged_row_passes_equal_in_place(precision):
    """The LDS-staged row kernels (launch variant 1) do the same arithmetic as
    the in-place four-step ones they were written against (variant 2): a
    7-step run with every row pass staged is bit-equal to one with every row
    pass in place.  (The specialised default kernels, variant 0, may contract
    multiply-adds differently; tests/test_gpu_variants.py bounds that.)"""
    import lesgo_amd
    from lesgo_amd.case import tune_clear, tune_get, tune_set
    cfg = cfg_of(48, 32, 16, precision=precision)
    passes = (("r2c", 48, 32), ("c2r", 48, 32), ("r2c", 72, 48), ("c2r", 72, 48))
    runs = {}
    for v in (2, 1):
        tune_clear()
        for kind, N, other in passes:
            n1, n2, S, _ = tune_get(kind, precision, N, other)
            assert n1 > 0
            tune_set(kind, precision, N, other, n1, n2, S, v)
            assert tune_get(kind, precision, N, other)[3] == v
        runs[v], _ = alone(cfg, 3, 1e-4, 1.0, True, 7)
    tune_clear()
    ref, got = runs[2], runs[1]
    for n in FIELDS_CMP:
        d = np.max(np.abs(got[n] - ref[n]))
        assert d == 0.0, (precision, n, d)
