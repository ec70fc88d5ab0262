"""The channel, ensemble and serving workloads at their own sizes.

Channel at 128^3 (Lx = 2 pi, Ly = pi, Lz = 1, dt = 2e-4): the FP64 step
against the oracle through a fresh start's Euler, Smagorinsky and first
dynamic steps at 1e-10, and the mixed step against the FP64 step for seeds
0-3 after 100 steps on the device (lg_diff_l2), velocity within 1e-4.

Ensemble of 48 members at 96 x 96 x 64 (seeds 0-47, z0 log-uniform in
[1e-5, 1e-3], mean pressure gradient uniform in [0.5, 1.5], odd members with
a 2 x 2 disk array, D = 0.15): members of the mixed batch equal the same
cases run alone bit for bit after 10 steps; a member of the FP64 batch
matches the oracle at 1e-10.

Serving on the grids {64^3, 96 x 48 x 64, 128 x 64 x 96, 128^3}: jobs
submitted while others run, mixed precision, each served result equal to
the job run alone bit for bit; an FP64 job served among them matches the
oracle at 1e-10.  Short horizons (DYN_init = 5 or 10) put the dynamic
model inside each run."""
import numpy as np
import pytest

from oracle import lesgo_oracle as o
from gpu_util import gpu_case, rel_l2, upload_all

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


# ---------------------------------------------------------------- channel 128^3
def test_channel128_fp64_matches_oracle():
    p = o.Params(128, 128, 128, initu=False, dyn_init=5)
    sp = o.Spectral(p)
    s = o.initial_condition(p, 0)
    with gpu_case(p) as c:
        upload_all(c, s)
        for chunk in (1, 4):
            for _ in range(chunk):
                o.step(sp, s)
            c.step(chunk)
            X = np.s_[1:p.nz, :, :p.nx]
            for n in ("u", "v", "w", "p", "Cs_opt2", "F_LM"):
                e = rel_l2(c.download(n)[X], s.f[n][X])
                print(f"  128^3 step {s.jt_total} {n}: {e:.2e}", flush=True)
                assert e <= 1e-10, (s.jt_total, n, e)


def test_channel128_mixed_vs_fp64_seeds_0_3():
    import lesgo_amd
    seeds = (0, 1, 2, 3)
    cases = {}
    for prec in ("fp64", "mixed"):
        cfg = lesgo_amd.Config(128, 128, 128, precision=prec, initu=False)
        c = lesgo_amd.Case(cfg, members=len(seeds))
        for m, sd in enumerate(seeds):
            c.seed_state(sd, 0.1, member=m)
        c.set_counters(0, 0)
        cases[prec] = c
    try:
        for prec in cases:
            cases[prec].step_graph(100)
        for m, sd in enumerate(seeds):
            e = {f: cases["mixed"].diff_l2(cases["fp64"], f, 1, 128 if f != "w" else 129,
                                           member=m, other_member=m) for f in ("u", "v", "w")}
            print(f"  seed {sd}: " + " ".join(f"{k}={v:.2e}" for k, v in e.items()), flush=True)
            for f, v in e.items():
                assert v <= 1e-4, (sd, f, v)
    finally:
        for c in cases.values():
            c.close()


# ------------------------------------------------------- ensemble 48 x 96x96x64
E = dict(members=48, nx=96, ny=96, Nz=64, dt=2e-4, dia=0.15, hub=0.1, ct=1.33)


def ensemble_members(n, rng_seed=0):
    rng = np.random.default_rng(rng_seed)
    zo = 10.0 ** rng.uniform(-5.0, -3.0, n)
    mpf = rng.uniform(0.5, 1.5, n)
    return [(m, float(zo[m]), float(mpf[m]), m % 2 == 1) for m in range(n)]


def _ens_cfg(precision):
    import lesgo_amd
    return lesgo_amd.Config(E["nx"], E["ny"], E["Nz"], dt=E["dt"], precision=precision,
                            initu=False, dyn_init=5)


def _ens_farm(cfg):
    import lesgo_amd
    return lesgo_amd.aligned_farm(cfg, 2, 2, dia=E["dia"], hub=E["hub"], sx_d=7.0, sy_d=5.0,
                                  ct_prime=E["ct"])


def _setup(c, m, cfg, seed, zo, mpf, turb):
    c.set_member(m, zo, mpf, 0.0)
    c.seed_state(seed, 0.1, member=m)
    if turb:
        c.turbines_init(_ens_farm(cfg), member=m)


def test_ensemble48_members_equal_alone():
    import lesgo_amd
    members = ensemble_members(E["members"])
    cfg = _ens_cfg("mixed")
    nsteps = 10
    check = (0, 1, 17, 46, 47)
    with lesgo_amd.Case(cfg, members=len(members)) as c:
        for m, mem in enumerate(members):
            _setup(c, m, cfg, *mem)
        c.set_counters(0, 0)
        c.step_graph(nsteps)
        got = {m: {n: c.download(n, member=m) for n in ("u", "v", "w", "p", "Cs_opt2")}
               for m in check}
    for m in check:
        with lesgo_amd.Case(cfg) as a:
            _setup(a, 0, cfg, *members[m])
            a.set_counters(0, 0)
            a.step_graph(nsteps)
            for n, g in got[m].items():
                d = np.max(np.abs(g - a.download(n)))
                print(f"  member {m} {n}: max |batch - alone| = {d:.1e}", flush=True)
                assert d == 0.0, (m, n, d)


def test_ensemble48_fp64_member_matches_oracle():
    import lesgo_amd
    members = ensemble_members(E["members"])
    cfg = _ens_cfg("fp64")
    m = 1                                  # odd: carries disks
    seed, zo, mpf, turb = members[m]
    p = o.Params(E["nx"], E["ny"], E["Nz"], dt=E["dt"], zo=zo, mean_p_force_x=mpf,
                 initu=False, dyn_init=5)
    sp = o.Spectral(p)
    s = o.initial_condition(p, seed)
    s.turbines = o.Turbines(p, [o.TurbineSpec(*t) for t in _ens_farm(cfg)])
    nsteps = 10
    with lesgo_amd.Case(cfg, members=len(members)) as c:
        for k, mem in enumerate(members):
            _setup(c, k, cfg, *mem)
        c.set_counters(0, 0)
        c.step_graph(nsteps)
        for _ in range(nsteps):
            o.step(sp, s)
        X = np.s_[1:p.nz, :, :p.nx]
        for n in ("u", "v", "w", "p", "Cs_opt2"):
            e = rel_l2(c.download(n, member=m)[X], s.f[n][X])
            print(f"  member {m} {n}: {e:.2e}", flush=True)
            assert e <= 1e-10, (n, e)
        udt, _ = c.turbines_state(member=m)
        np.testing.assert_allclose(udt, s.turbines.u_d_T, rtol=1e-10)


# ------------------------------------------------------------------- serving
GRIDS = ((64, 64, 64), (96, 48, 64), (128, 64, 96), (128, 128, 128))
# grid index, seed, z0, steps, disks, submit at tick
JOBS = [
    (0, 11, 1e-4, 14, False, 0),
    (3, 12, 3e-5, 12, True, 0),
    (1, 13, 5e-4, 16, False, 1),
    (2, 14, 2e-5, 13, True, 3),
    (0, 15, 1e-3, 9, True, 4),
    (3, 16, 7e-5, 10, False, 6),
    (2, 17, 2e-4, 11, False, 9),
    (1, 18, 1e-5, 15, True, 12),
]


def _job_cfg(grid, zo, precision="mixed"):
    import lesgo_amd
    return lesgo_amd.Config(*grid, zo=zo, precision=precision, initu=False, dyn_init=10)


def _job_disks(cfg, on):
    import lesgo_amd
    return lesgo_amd.aligned_farm(cfg, 2, 2, dia=0.15, hub=0.1, sx_d=7.0, sy_d=5.0,
                                  ct_prime=1.33) if on else []


def test_serving_grids_equal_alone_and_oracle():
    import lesgo_amd
    from lesgo_amd.serve import Server
    ids = {}
    fp64_job = (0, 21, 3e-4, 12)           # a 64^3 FP64 job among the mixed ones
    with Server(100e9, 16) as srv:
        tick = 0
        while True:
            for k, (g, seed, zo, n, disks, at) in enumerate(JOBS):
                if at == tick:
                    cfg = _job_cfg(GRIDS[g], zo)
                    ids[k] = srv.submit(cfg, seed, n, _job_disks(cfg, disks), keep=True)
            if tick == 2:
                g, seed, zo, n = fp64_job
                ids["fp64"] = srv.submit(_job_cfg(GRIDS[g], zo, "fp64"), seed, n, keep=True)
            _, running, queued = srv.tick()
            tick += 1
            if tick > 12 and running == 0 and queued == 0:
                break
            assert tick < 500, "serving loop did not drain"
        res = {}
        for k, i in ids.items():
            info = srv.job(i)
            assert info["state"] == 2, (k, info)
            g = GRIDS[JOBS[k][0]] if k != "fp64" else GRIDS[fp64_job[0]]
            shape = (g[2] + 2, g[1], 2 * (g[0] // 2 + 1))
            res[k] = {n: srv.result(i, n, shape) for n in ("u", "v", "w", "p")}
    print(f"  drained in {tick} ticks", flush=True)
    for k, (g, seed, zo, n, disks, at) in enumerate(JOBS):
        cfg = _job_cfg(GRIDS[g], zo)
        with lesgo_amd.Case(cfg) as a:
            a.seed_state(seed)
            if disks:
                a.turbines_init(_job_disks(cfg, True))
            a.step_graph(n)
            for f in ("u", "v", "w", "p"):
                d = np.max(np.abs(res[k][f] - a.download(f)))
                assert d == 0.0, (k, GRIDS[g], f, d)
        print(f"  job {k} {GRIDS[g]}: served == alone", flush=True)
    g, seed, zo, n = fp64_job
    p = o.Params(*GRIDS[g], zo=zo, initu=False, dyn_init=10)
    sp = o.Spectral(p)
    s = o.initial_condition(p, seed)
    for _ in range(n):
        o.step(sp, s)
    for f, lv in (("u", np.s_[1:p.nz]), ("v", np.s_[1:p.nz]), ("w", np.s_[1:p.nz + 1]),
                  ("p", np.s_[0:p.nz + 1])):
        e = rel_l2(res["fp64"][f][lv][:, :, :p.nx], s.f[f][lv][:, :, :p.nx])
        print(f"  served FP64 job {f}: {e:.2e}", flush=True)
        assert e <= 1e-10, (f, e)
