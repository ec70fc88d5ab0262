"""Every transform length the benchmark configurations run, against the
oracle, through the subsystem entry points and a short whole step.

The x-y transforms are instantiated per length: four-step register kernels
for the row passes (r2c, c2r and the Lagrangian model's fused row passes)
and the column passes (forward, inverse with multipliers, the 3/2-grid
padding and unpadding maps).  The grids below put each length the
configurations use in each of those kernels:

  256 x 256: rows and columns of 256, the 3/2 grid's 384 (wind farm)
  128 x 128: 128 and 192 (channel and serving grids)
  96 x 96:   96 and 144 (ensemble members)
  96 x 48:   48, 72 (columns), 96, 144 (rows) (serving grid)
  128 x 64:  64, 96 (columns), 128, 192 (rows) (serving grid)
  64 x 64:   64 and 96 (serving grid)

Each grid runs filt_da, divstress_uv/w, convec, press_stag_array + project
and sgs_stag with the Lagrangian model (filter_cutoff bands at 2 and 4
delta) in FP64 at 1e-11 (1e-10 for the dynamic model) relative L2 and in
mixed precision at the bars below, and a six-step FP64 fresh start (Euler
step, Smagorinsky phase, F_* initialisation and one dynamic step, the
spectral-resident step's own fused passes) at 1e-10."""
import numpy as np
import pytest

from oracle import lesgo_oracle as o
from gpu_util import gpu_case, oracle_state, rel_l2, upload_all

pytestmark = pytest.mark.gpu

GRIDS = [(256, 256), (128, 128), (96, 96), (96, 48), (128, 64), (64, 64)]
NZ = 8
# mixed precision: inputs are rounded to float on upload (6e-8) and every
# transform runs in float; the dynamic model's Cs_opt2 divides differences
# of nearly equal filtered products, which magnifies the rounding ~100-fold
TOL = {"fp64": 1e-11, "mixed": 2e-5}
TOL_DYN = {"fp64": 1e-10, "mixed": 2e-3}

_BASES = {}


def base(nx, ny):
    if (nx, ny) not in _BASES:
        _BASES.clear()
        _BASES[(nx, ny)] = oracle_state(nx, ny, NZ, steps=3)
    return _BASES[(nx, ny)]


def check(c, s, name, levels, tol, cols=None):
    p = s.p
    g = c.download(name)[levels, :, : (cols or p.nx)]
    r = s.f[name][levels, :, : (cols or p.nx)]
    err = rel_l2(g, r)
    print(f"  {p.nx}x{p.ny} {name}: {err:.2e}", flush=True)
    assert err <= tol, f"{p.nx}x{p.ny} {name}: rel L2 {err:.3e} > {tol:.1e}"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nx,ny", GRIDS)
def test_subsystems_at_length(nx, ny, precision):
    p, sp, s0 = base(nx, ny)
    tol = TOL[precision]
    nz = p.nz
    # filt_da: forward rows and columns, inverse columns with i kx, i ky
    s = s0.copy()
    with gpu_case(p, precision) as c:
        upload_all(c, s)
        c.filt_da("u", "dudx", "dudy")
        s.u[:], s.dudx[:], s.dudy[:] = o.filt_da(sp, s.u)
        for n in ("u", "dudx", "dudy"):
            check(c, s, n, slice(1, nz + 1), tol)
    # divstress_uv/w: ddx, ddy, ddxy
    s = s0.copy()
    with gpu_case(p, precision) as c:
        upload_all(c, s)
        c.divstress_uv()
        c.divstress_w()
        o.divstress_uv(sp, s)
        o.divstress_w(sp, s)
        check(c, s, "divtx", slice(1, nz), tol, cols=p.ld)
        check(c, s, "divty", slice(1, nz), tol, cols=p.ld)
        check(c, s, "divtz", slice(1, nz + 1), tol)
    # convec: padding to the 3/2 grid, its row passes, unpadding
    s = s0.copy()
    with gpu_case(p, precision) as c:
        upload_all(c, s)
        c.convec()
        o.convec(sp, s)
        check(c, s, "RHSx", slice(1, nz), tol)
        check(c, s, "RHSy", slice(1, nz), tol)
        check(c, s, "RHSz", slice(1, nz + 1), tol)
    # press_stag_array + project on a non-solenoidal u*
    s = s0.copy()
    rng = np.random.default_rng(12)
    for n in ("u", "v", "w"):
        s.f[n][1:nz, :, :nx] += 0.5 * rng.standard_normal((nz - 1, ny, nx))
    with gpu_case(p, precision) as c:
        upload_all(c, s)
        c.press_stag_array()
        o.press_stag_array(sp, s)
        check(c, s, "p", slice(0, nz + 1), tol)
        check(c, s, "dpdx", slice(1, nz), tol)
        check(c, s, "dpdz", slice(1, nz + 1), tol)
        c.project()
        o.project(p, s)
        for n in ("u", "v", "w"):
            check(c, s, n, slice(1, nz + 1), tol)
    # sgs_stag with the Lagrangian model (jt_total = cs_count)
    s = s0.copy()
    s.jt_total = 5
    with gpu_case(p, precision) as c:
        upload_all(c, s)
        c.sgs_stag()
        o.sgs_stag(sp, s)
        for n in ("F_LM", "F_MM", "F_QN", "F_NN", "Cs_opt2", "txx", "txy", "tzz"):
            check(c, s, n, slice(1, nz), TOL_DYN[precision])


@pytest.mark.parametrize("nx,ny", GRIDS)
def test_fresh_start_steps_at_length(nx, ny):
    """Six steps of a fresh start with DYN_init = 5 (Euler step 1,
    Smagorinsky steps 2-4, F_* initialised at step 5, dynamic step 5 and
    the spectral-resident step's fused passes) against the oracle at 1e-10
    on u, v, w, p, Cs_opt2 and F_LM; rmsdiv (from the start-of-step
    derivatives: the seeded field's divergence at step 1, roundoff after a
    projection later) at 1e-10 relative or 1e-12 absolute."""
    p = o.Params(nx, ny, NZ, initu=False, dyn_init=5)
    sp = o.Spectral(p)
    s = o.initial_condition(p, 4)
    with gpu_case(p) as c:
        upload_all(c, s)
        for chunk in (1, 4, 1):
            for _ in range(chunk):
                o.step(sp, s)
            c.step(chunk)
            X = np.s_[1:p.nz, :, :p.nx]
            for n in ("u", "v", "w", "p", "Cs_opt2", "F_LM"):
                e = rel_l2(c.download(n)[X], s.f[n][X])
                print(f"  {nx}x{ny} step {s.jt_total} {n}: {e:.2e}", flush=True)
                assert e <= 1e-10, (s.jt_total, n, e)
            rg, rr = c.rmsdiv(), o.rmsdiv(p, s)
            print(f"  {nx}x{ny} step {s.jt_total} rmsdiv gpu {rg:.3e} oracle {rr:.3e}")
            assert abs(rg - rr) <= 1e-10 * rr + 1e-12
