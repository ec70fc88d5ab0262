"""The convective products formed in the 3/2 grid's forward row pass
(k_rows_r2c_prod) against the separate passes (LESGO_AMD_PROD_ROWS=0:
k_products writes the three product fields, the row pass reads them).  The
fused pass forms the products with k_products' expressions in FP64 and
transforms them with k_rows_r2c_v4's four-step arithmetic, so the flow
agrees to rounding: the transforms' multiply-adds may be contracted in
another order in the other kernel (bound: FP64 1e-13 relative L2 after 15
steps, mixed 1e-5 on u, v, w, p and 1e-4 elsewhere, as for the launch
variants).  A grid whose 3/2 length has no four-step split takes the
separate passes either way, bit for bit.  Both paths match the oracle
(tests/test_gpu_lengths.py and the other parity tests run the fused one)."""
import os

import numpy as np
import pytest

from gpu_util import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "RHSx", "RHSy", "RHSz", "Cs_opt2")


def _case(fused, nx, ny, nz, precision, members):
    import lesgo_amd
    if not fused:
        os.environ["LESGO_AMD_PROD_ROWS"] = "0"
    try:
        cfg = lesgo_amd.Config(nx, ny, nz, initu=False, dyn_init=6, precision=precision)
        return lesgo_amd.Case(cfg, members=members)
    finally:
        os.environ.pop("LESGO_AMD_PROD_ROWS", None)


def _run_pair(nx, ny, nz, precision, tol, members=1, farm=False):
    """tol(field): largest allowed relative L2 gap (0: the same bits)."""
    import lesgo_amd
    cases = [_case(False, nx, ny, nz, precision, members),
             _case(True, nx, ny, nz, precision, members)]
    K = np.s_[1:nz]
    try:
        for c in cases:
            for m in range(members):
                c.seed_state(3 + m, member=m) if members > 1 else c.seed_state(3)
            if farm:
                c.turbines_init(lesgo_amd.aligned_farm(c.cfg, 2, 2, dia=0.3, hub=0.3,
                                                       sx_d=7.0, sy_d=5.0))
        for chunk in (1, 9, 5):
            for c in cases:
                c.step_graph(chunk)
            step = cases[0].counters()[1]
            worst = 0.0
            for m in range(members):
                for n in FIELDS:
                    a = cases[0].download(n, member=m)[K][:, :, :nx]
                    b = cases[1].download(n, member=m)[K][:, :, :nx]
                    e = rel_l2(b, a)
                    worst = max(worst, e)
                    assert e <= tol(n), (n, m, step, e)
            print(f"  {nx}x{ny}x{nz} {precision} x{members} step {step}: worst {worst:.2e}",
                  flush=True)
    finally:
        for c in cases:
            c.close()


def _tol(precision):
    if precision == "fp64":
        return lambda f: 1e-13
    return lambda f: 1e-5 if f in ("u", "v", "w", "p") else 1e-4


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_products_equal_separate_passes(precision):
    # 48 x 32 x 16: nx2 = 72 (split 8 x 9); Euler step, Smagorinsky steps,
    # dynamic calls at steps 6, 10 and 15
    _run_pair(48, 32, 16, precision, _tol(precision))


def test_fused_products_batch_and_384_rows():
    # the 3/2 grid of the bench's x length (nx2 = 384, split 16 x 24), two
    # members and a disk array
    _run_pair(256, 32, 8, "fp64", _tol("fp64"), members=2, farm=True)


def test_grid_without_fused_products_takes_separate_passes():
    # nx = 36: nx2 = 54 has no four-step split; both cases run k_products
    _run_pair(36, 32, 12, "fp64", lambda f: 0.0)
