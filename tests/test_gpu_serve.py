"""Serving loop: jobs of two grid classes arrive while others run, share
pools of two slots (so pools are created, filled, compacted and released)
and each job's final u, v, w, p equal, bit for bit, the same job run alone
(same kernels, launch choices and per-member arithmetic; joins happen only
at dynamic-model boundaries)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# grid, seed, zo, nsteps, disks, submit at tick
JOBS = [
    ((32, 32, 16), 1, 1e-4, 11, False, 0),
    ((48, 32, 16), 2, 3e-5, 17, True, 0),
    ((32, 32, 16), 3, 5e-4, 8, True, 2),
    ((32, 32, 16), 4, 2e-5, 13, False, 3),
    ((48, 32, 16), 5, 1e-3, 9, False, 6),
    ((32, 32, 16), 6, 7e-5, 6, False, 7),
]


def cfg_of(grid, zo, fresh=False):
    """fresh: the fresh-start schedule with DYN_init = 10, so jobs admitted
    at different ticks sit in one pool at different points of it (first
    Euler step, Smagorinsky phase, F_* initialisation, dynamic)."""
    import lesgo_amd
    return lesgo_amd.Config(*grid, zo=zo, precision="mixed", initu=not fresh, dyn_init=10)


def disks_of(cfg, on):
    import lesgo_amd
    if not on:
        return []
    return lesgo_amd.aligned_farm(cfg, 2, 2, dia=0.3, hub=0.3, sx_d=7, sy_d=5)


def alone(grid, seed, zo, nsteps, disks, fresh=False):
    import lesgo_amd
    cfg = cfg_of(grid, zo, fresh)
    with lesgo_amd.Case(cfg) as c:
        c.seed_state(seed)
        if disks:
            c.turbines_init(disks_of(cfg, True))
        c.step_graph(nsteps)
        return {n: c.download(n) for n in ("u", "v", "w", "p")}


def run_server(budget, pool_members=2, max_ticks=400, fresh=False):
    from lesgo_amd.serve import Server
    ids = {}
    with Server(budget, pool_members) as s:
        tick = 0
        while True:
            for k, (grid, seed, zo, nsteps, disks, at) in enumerate(JOBS):
                if at == tick:
                    cfg = cfg_of(grid, zo, fresh)
                    ids[k] = s.submit(cfg, seed, nsteps, disks_of(cfg, disks), keep=True)
            _, running, queued = s.tick()
            tick += 1
            st = s.stats()
            assert st["bytes_used"] <= budget
            if tick > max(j[5] for j in JOBS) and running == 0 and queued == 0:
                break
            assert tick < max_ticks, "serving loop did not drain"
        out = {}
        for k, (grid, seed, zo, nsteps, disks, at) in enumerate(JOBS):
            info = s.job(ids[k])
            assert info["state"] == 2 and info["steps_done"] == nsteps, info
            assert info["t_arrive"] <= info["t_admit"] <= info["t_done"]
            shape = (grid[2] + 2, grid[1], 2 * (grid[0] // 2 + 1))
            out[k] = {n: s.result(ids[k], n, shape) for n in ("u", "v", "w", "p")}
        return out, tick


def check_alone(out, fresh=False):
    for k, (grid, seed, zo, nsteps, disks, at) in enumerate(JOBS):
        ref = alone(grid, seed, zo, nsteps, disks, fresh)
        for n in ("u", "v", "w", "p"):
            d = np.max(np.abs(out[k][n] - ref[n]))
            print(f"  job {k} {grid} {n}: max |served - alone| = {d:.3e}", flush=True)
            assert d == 0.0, (k, n, d)


@pytest.mark.parametrize("fresh", [False, True])
def test_served_jobs_equal_alone(fresh):
    out, ticks = run_server(budget=50e9, fresh=fresh)
    print(f"  drained in {ticks} ticks", flush=True)
    check_alone(out, fresh)


def test_budget_limits_admission():
    """Budget for one two-slot pool of the larger grid only: the grid classes
    take turns (an empty pool is released for the other class) and every
    job still finishes with the result it has alone."""
    import ctypes as C
    import lesgo_amd
    from lesgo_amd import _lib
    b = C.c_double()
    _lib.check(_lib.lib().lg_case_bytes(C.byref(cfg_of((48, 32, 16), 1e-4).to_c()), 2,
                                        C.byref(b)), "case_bytes")
    out, ticks = run_server(budget=1.2 * b.value)
    print(f"  budget {1.2 * b.value / 1e6:.1f} MB: drained in {ticks} ticks", flush=True)
    check_alone(out)
    assert lesgo_amd  # imported for the library load


def test_case_bytes_matches_device_allocation():
    import ctypes as C
    import lesgo_amd
    from lesgo_amd import _lib
    cfg = lesgo_amd.Config(96, 96, 64, precision="mixed")
    b = C.c_double()
    _lib.check(_lib.lib().lg_case_bytes(C.byref(cfg.to_c()), 4, C.byref(b)), "case_bytes")

    def free():
        f, t = C.c_double(), C.c_double()
        _lib.check(_lib.lib().lg_mem_info(C.byref(f), C.byref(t)), "mem_info")
        return f.value

    c0 = lesgo_amd.Case(lesgo_amd.Config(32, 32, 8))       # runtime set up before measuring
    free0 = free()
    c = lesgo_amd.Case(cfg, members=4)
    free1 = free()
    c.close()
    c0.close()
    used = free0 - free1
    print(f"  case_bytes {b.value / 1e6:.1f} MB, device free dropped {used / 1e6:.1f} MB",
          flush=True)
    assert 0.95 * b.value <= used <= 1.1 * b.value + 64e6
