"""Shared helpers for the GPU parity tests (oracle vs the HIP path via the C ABI)."""
import numpy as np

from oracle import lesgo_oracle as o


def oracle_state(nx=32, ny=32, Nz=16, steps=3, seed=0, turbines=None, **kw):
    p = o.Params(nx, ny, Nz, **kw)
    sp = o.Spectral(p)
    s = o.initial_condition(p, seed)
    if turbines is not None:
        s.turbines = o.Turbines(p, turbines)
    for _ in range(steps):
        o.step(sp, s)
    return p, sp, s


def gpu_case(p, precision="fp64"):
    import lesgo_amd
    cfg = lesgo_amd.Config(p.nx, p.ny, p.Nz, lx=p.lx, ly=p.ly, lz=p.lz, dt=p.dt, zo=p.zo,
                           vonk=p.vonk, cs_count=p.cs_count,
                           use_mean_p_force=p.use_mean_p_force,
                           mean_p_force_x=p.mean_p_force_x,
                           mean_p_force_y=p.mean_p_force_y, nu=p.nu, precision=precision,
                           initu=p.initu, dyn_init=p.dyn_init, use_cfl_dt=p.use_cfl_dt,
                           cfl=p.cfl)
    return lesgo_amd.Case(cfg)


def upload_all(c, s):
    for n, a in s.f.items():
        c.upload(n, a)
    c.set_counters(s.jt, s.jt_total)


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def assert_close(c, s, name, levels, tol, cols=None):
    p = s.p
    g = c.download(name)[levels, :, : (cols or p.nx)]
    r = s.f[name][levels, :, : (cols or p.nx)]
    err = rel_l2(g, r)
    print(f"  {name}: rel L2 = {err:.3e}", flush=True)
    assert err <= tol, f"{name}: rel L2 {err:.3e} > {tol:.1e}"
    return err
