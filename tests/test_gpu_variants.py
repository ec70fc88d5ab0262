"""Launch variants of one pass compute the same bits.

The inverse column pass has a kernel specialised by index map and spectral
multiplier (variant 0, the default) and the unspecialised four-step kernel
(variant 2); both do the same arithmetic.  A 256 x 256 x 24 wind-farm case
(disks, fresh start: Euler step, Smagorinsky steps, F_* initialisation and a
Lagrangian update, so every multiplier, the test-filter bands and the 3/2-grid
padding run) is stepped with each variant on every split of the 256-point
(filt_da, divstress, test filters) and 384-point (convec's padding) inverse
columns.  With the (24, 16) split of the padded pass u, v, w, p, RHS, Cs_opt2,
F_LM and the stresses are equal bit for bit.  With (16, 24) the specialised
kernel knows which inputs are the padding's zeros at compile time, the
compiler folds those butterflies and contracts the remaining multiply-adds
into FMAs in another order, so the results differ by rounding: FP64 within
1e-12 relative L2 after 7 steps; mixed within 1e-5 on u, v, w, p and 1e-4 on
the rest (float rounding grown over the steps; the dynamic coefficient, a
ratio of filtered products, moves most: 2e-5).  Both variants match the oracle (tests/test_gpu_lengths.py,
tests/test_gpu_windfarm256.py); a case gets the same variant alone or in a
batch (the choice is keyed by the grid), so batch equality is unaffected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "RHSx", "RHSz_f", "Cs_opt2", "F_LM", "F_QN", "txz", "divtx",
          "dudx", "dwdy")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("pad_split", [(16, 24), (24, 16)])
def test_inverse_column_variants_equal(precision, pad_split):
    import lesgo_amd
    from lesgo_amd.case import tune_clear, tune_get, tune_set
    cfg = lesgo_amd.Config(256, 256, 24, precision=precision, initu=False, dyn_init=5)
    farm = lesgo_amd.aligned_farm(cfg, 2, 2, dia=0.3, hub=0.3, sx_d=7.0, sy_d=5.0)
    out = {}
    try:
        for v in (2, 0):
            tune_clear()
            tune_set("cols_inv", precision, 256, 128, 16, 16, 8, v)
            tune_set("cols_pad", precision, 384, 128, pad_split[0], pad_split[1], 8, v)
            assert tune_get("cols_pad", precision, 384, 128)[3] == v
            with lesgo_amd.Case(cfg) as c:
                c.seed_state(5)
                c.turbines_init(farm)
                c.step_graph(7)
                out[v] = {f: c.download(f) for f in FIELDS}
    finally:
        tune_clear()
    for f in FIELDS:
        tol = 1e-12 if precision == "fp64" else (1e-5 if f in ("u", "v", "w", "p") else 1e-4)
        a, b = out[2][f][1:24, :, :256], out[0][f][1:24, :, :256]
        d = float(np.max(np.abs(a - b)))
        e = float(np.linalg.norm(a - b) / max(np.linalg.norm(a), 1e-300))
        print(f"  {precision} pad {pad_split} {f}: max |v4 - v2| = {d:.1e}, rel L2 {e:.1e}",
              flush=True)
        if pad_split == (24, 16):
            assert np.array_equal(a, b), (f, d)
        else:
            assert e <= tol, (f, e)
