"""Launch variants of one pass compute the same results.

The column and row passes each have a kernel specialised at compile time
(by index map and multiplier for columns, by stored band for inverse rows,
32-bit offsets for forward rows; variant 0, the default, and variant 3, the same with the twiddles read
from global memory instead of an LDS copy) and the unspecialised four-step
kernel (variant 2); all do the same arithmetic.  A 256 x 256 x 24 wind-farm case
(disks, fresh start: Euler step, Smagorinsky steps, F_* initialisation and a
Lagrangian update, so every multiplier, the test-filter bands, the x
derivative's second output and the 3/2-grid padding run) is stepped with
each variant of one pass.

Where the specialised kernel sees the same arithmetic graph the results are
equal bit for bit: the 256-point inverse columns and rows, the (24, 16)
split of the padded columns, the forward columns (plain and unpadding), the
forward rows, and variants 3 and 0 everywhere.  Where it knows at compile time that some
inputs are the padding's zeros ((16, 24) padded columns, the 384-point rows
whose input band is a third of the transform), the compiler folds those
butterflies and contracts the remaining multiply-adds into FMAs in another
order, so the results differ by rounding: FP64 within 1e-12 relative L2
after 7 steps; mixed within 1e-5 on u, v, w, p and 1e-4 on the rest (float
rounding grown over the steps; the dynamic coefficient, a ratio of filtered
products, moves most).  Every variant matches the oracle
(tests/test_gpu_lengths.py, tests/test_gpu_windfarm256.py); a case gets the
same variant alone or in a batch (the choice is keyed by the grid), so batch
equality is unaffected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "RHSx", "RHSz_f", "Cs_opt2", "F_LM", "F_QN", "txz", "divtx",
          "dudx", "dwdy")

# name: ([(pass kind, N, other, n1, n2, per_wg)], (variant a, variant b), bit-equal)
CASES = {
    "cols": ([("cols_inv", 256, 128, 16, 16, 8), ("cols_pad", 384, 128, 24, 16, 8)], (2, 0),
             True),
    "cols_pad16x24": ([("cols_pad", 384, 128, 16, 24, 8)], (2, 0), False),
    "rows256": ([("c2r", 256, 256, 16, 16, 16)], (2, 0), True),
    "rows384": ([("c2r", 384, 384, 16, 24, 8)], (2, 0), False),
    "cols_fwd": ([("cols_fwd", 256, 128, 16, 16, 8), ("cols_unpad", 384, 128, 16, 24, 8)],
                 (2, 0), True),
    "rows_fwd": ([("r2c", 256, 256, 16, 16, 12), ("r2c", 384, 384, 16, 24, 8)], (2, 0), True),
    # twiddles from global memory (3) or the block's LDS copy (0): the same values
    "twiddles": ([("cols_inv", 256, 128, 16, 16, 8), ("cols_pad", 384, 128, 16, 24, 8),
                  ("c2r", 256, 256, 16, 16, 16), ("c2r", 384, 384, 16, 24, 8),
                  ("cols_fwd", 256, 128, 16, 16, 8), ("cols_unpad", 384, 128, 16, 24, 8),
                  ("r2c", 256, 256, 16, 16, 12), ("r2c", 384, 384, 16, 24, 8)],
                 (3, 0), True),
}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("which", sorted(CASES))
def test_inverse_pass_variants_agree(precision, which):
    import lesgo_amd
    from lesgo_amd.case import tune_clear, tune_get, tune_set
    cfg = lesgo_amd.Config(256, 256, 24, precision=precision, initu=False, dyn_init=5)
    farm = lesgo_amd.aligned_farm(cfg, 2, 2, dia=0.3, hub=0.3, sx_d=7.0, sy_d=5.0)
    out = {}
    passes, variants, exact = CASES[which]
    try:
        for v in variants:
            tune_clear()
            for kind, N, other, n1, n2, S in passes:
                tune_set(kind, precision, N, other, n1, n2, S, v)
                assert tune_get(kind, precision, N, other)[3] == v
            with lesgo_amd.Case(cfg) as c:
                c.seed_state(5)
                c.turbines_init(farm)
                c.step_graph(7)
                out[v] = {f: c.download(f) for f in FIELDS}
    finally:
        tune_clear()
    for f in FIELDS:
        a, b = out[variants[0]][f][1:24, :, :256], out[variants[1]][f][1:24, :, :256]
        d = float(np.max(np.abs(a - b)))
        e = float(np.linalg.norm(a - b) / max(np.linalg.norm(a), 1e-300))
        print(f"  {precision} {which} {f}: max |variant {variants[1]} - {variants[0]}| = {d:.1e}, "
              f"rel L2 {e:.1e}", flush=True)
        if exact:
            assert np.array_equal(a, b), (f, d)
        else:
            tol = 1e-12 if precision == "fp64" else (1e-5 if f in ("u", "v", "w", "p") else 1e-4)
            assert e <= tol, (f, e)
