"""Running time averages (tavg) and streamwise spectra on the device against
the oracle's Stats, and the cross-case difference reduction."""
import numpy as np
import pytest

from oracle import lesgo_oracle as o
from gpu_util import gpu_case, rel_l2

pytestmark = pytest.mark.gpu


def test_tavg_and_spectra_match_oracle():
    """48x32x16 with a 2x2 disk array, 9 steps, sums from step 3 every
    step: all 26 time averages (levels 1..nz, x < nx) and the three spectra
    at 1e-10 relative L2 (or 1e-12 absolute for fields that vanish)."""
    import lesgo_amd
    p = o.Params(48, 32, 16)
    specs = o.aligned_farm(p, 2, 2, dia=0.3, hub=0.3, sx_d=7, sy_d=5)
    sp = o.Spectral(p)
    s = o.initial_condition(p, 5)
    s.turbines = o.Turbines(p, specs)
    st = o.Stats(p, 2, 9, 1)
    with gpu_case(p) as c:
        for n, a in s.f.items():
            c.upload(n, a)
        c.set_counters(0, 0)
        c.turbines_init([(t.xloc, t.yloc, t.height, t.dia, t.thk, t.theta1, t.theta2, t.ct_prime)
                         for t in specs])
        c.stats_enable(2, 9, 1)
        for _ in range(9):
            o.step(sp, s)
            st.after_step(sp, s)
        c.step(9)
        tt, init = c.stats_info()
        assert init and abs(tt - st.total_time) < 1e-15
        K = np.s_[1:p.nz + 1]
        for n in o.TAVG_NAMES:
            g, r = c.stats_mean(n)[K], st.mean(n)[K]
            nr = np.linalg.norm(r)
            e = np.linalg.norm(g - r) / nr if nr > 1e-10 else np.abs(g - r).max()
            print(f"  {n}: {e:.2e}", flush=True)
            assert e <= (1e-10 if nr > 1e-10 else 1e-12), (n, e)
        for comp in range(3):
            e = rel_l2(c.stats_spectrum(comp)[1:p.nz], st.spectrum(comp)[1:p.nz])
            print(f"  spectrum {comp}: {e:.2e}", flush=True)
            assert e <= 1e-10, (comp, e)
        assert lesgo_amd


def test_diff_l2_across_precisions():
    import lesgo_amd
    cfg = lesgo_amd.Config(32, 32, 16)
    with lesgo_amd.Case(cfg) as a, lesgo_amd.Case(
            lesgo_amd.Config(32, 32, 16, precision="mixed")) as b:
        a.seed_state(1)
        b.seed_state(1)
        assert a.diff_l2(a, "u", 1, 16) == 0.0
        e0 = b.diff_l2(a, "u", 1, 16)
        assert 0.0 < e0 < 1e-6            # float rounding of the same start
        a.step(10)
        b.step(10)
        e = b.diff_l2(a, "u", 1, 16)
        u64, u32 = a.download("u")[1:17, :, :32], b.download("u")[1:17, :, :32]
        np.testing.assert_allclose(e, rel_l2(u32, u64), rtol=1e-9)
