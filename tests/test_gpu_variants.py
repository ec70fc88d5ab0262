"""Launch variants of one pass compute the same results.

The column and row passes each have a kernel specialised at compile time
(by index map and multiplier for columns, by stored band for inverse rows,
32-bit offsets for forward rows; variant 0, the default, and variant 3, the
same with the twiddles read from global memory instead of an LDS copy) and
the unspecialised four-step kernel (variant 2); all do the same arithmetic.  A 256 x 256 x 24 wind-farm case
(disks, fresh start: Euler step, Smagorinsky steps, F_* initialisation and a
Lagrangian update, so every multiplier, the test-filter bands, the x
derivative's second output and the 3/2-grid padding run) is stepped with
each variant of one pass.

The variants compute the same arithmetic, but where a specialised kernel
knows some inputs at compile time (the padding's zeros, a band that is a
third of the transform) or knows its lanes' index ranges, the compiler folds
those butterflies and may contract the remaining multiply-adds into FMAs in
another order, so results can differ by rounding (measured: bit-equal for
the twiddle source and for most 256-point passes; rel L2 1e-16 per step
where the contractions differ).  Bound: FP64 within 1e-12 relative L2 after
7 steps; mixed within 1e-5 on u, v, w, p and 1e-4 on the rest (float
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

# name: ([(pass kind, N, other, n1, n2, per_wg)], (variant a, variant b))
CASES = {
    "cols": ([("cols_inv", 256, 128, 16, 16, 8), ("cols_pad", 384, 128, 24, 16, 8)], (2, 0)),
    "cols_pad16x24": ([("cols_pad", 384, 128, 16, 24, 8)], (2, 0)),
    "rows256": ([("c2r", 256, 256, 16, 16, 16)], (2, 0)),
    "rows384": ([("c2r", 384, 384, 16, 24, 8)], (2, 0)),
    "cols_fwd": ([("cols_fwd", 256, 128, 16, 16, 8), ("cols_unpad", 384, 128, 16, 24, 8)],
                 (2, 0)),
    "rows_fwd": ([("r2c", 256, 256, 16, 16, 12), ("r2c", 384, 384, 16, 24, 8)], (2, 0)),
    # twiddles from global memory (3) or the block's LDS copy (0): the same values
    "twiddles": ([("cols_inv", 256, 128, 16, 16, 8), ("cols_pad", 384, 128, 16, 24, 8),
                  ("c2r", 256, 256, 16, 16, 16), ("c2r", 384, 384, 16, 24, 8),
                  ("cols_fwd", 256, 128, 16, 16, 8), ("cols_unpad", 384, 128, 16, 24, 8),
                  ("r2c", 256, 256, 16, 16, 12), ("r2c", 384, 384, 16, 24, 8)],
                 (3, 0)),
}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("which", sorted(CASES))
def test_inverse_pass_variants_agree(precision, which):
    import lesgo_amd
    from lesgo_amd.case import tune_clear, tune_get, tune_set
    cfg = lesgo_amd.Config(256, 256, 24, precision=precision, initu=False, dyn_init=5)
    farm = lesgo_amd.aligned_farm(cfg, 2, 2, dia=0.3, hub=0.3, sx_d=7.0, sy_d=5.0)
    out = {}
    passes, variants = CASES[which]
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
        tol = 1e-12 if precision == "fp64" else (1e-5 if f in ("u", "v", "w", "p") else 1e-4)
        assert e <= tol, (f, e)
