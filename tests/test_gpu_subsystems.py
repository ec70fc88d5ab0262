"""Per-subsystem parity of the HIP path against the oracle, called through the
C ABI on the same state (each subsystem replaces one reference routine)."""
import numpy as np
import pytest

from oracle import lesgo_oracle as o
from gpu_util import assert_close, gpu_case, oracle_state, upload_all

pytestmark = pytest.mark.gpu
TOL = 1e-11


@pytest.fixture(scope="module")
def base():
    return oracle_state(32, 24, 12, steps=3)


def test_filt_da(base):
    p, sp, s0 = base
    s = s0.copy()
    with gpu_case(p) as c:
        upload_all(c, s)
        c.filt_da("u", "dudx", "dudy")
        s.u[:], s.dudx[:], s.dudy[:] = o.filt_da(sp, s.u)
        for n in ("u", "dudx", "dudy"):
            assert_close(c, s, n, slice(1, p.nz + 1), TOL)


def test_ddz(base):
    p, sp, s0 = base
    s = s0.copy()
    with gpu_case(p) as c:
        upload_all(c, s)
        c.ddz_uv("v", "dvdz")
        c.ddz_w("w", "dwdz")
        s.dvdz[:] = o.ddz_uv(p, s.v, s.dvdz)
        s.dwdz[:] = o.ddz_w(p, s.w, s.dwdz)
        assert_close(c, s, "dvdz", slice(0, p.nz + 1), 1e-14)
        assert_close(c, s, "dwdz", slice(0, p.nz + 1), 1e-14)


def test_wallstress(base):
    p, sp, s0 = base
    s = s0.copy()
    with gpu_case(p) as c:
        upload_all(c, s)
        c.wallstress()
        o.wallstress(sp, s)
        for n in ("txz", "tyz", "dudz", "dvdz"):
            assert_close(c, s, n, slice(1, 2), TOL)
            assert_close(c, s, n, slice(p.nz, p.nz + 1), TOL)


@pytest.mark.parametrize("jt_total", [4, 5])
def test_sgs_stag(base, jt_total):
    """jt_total = 5 runs lagrange_Sdep (cs_count = 5), 4 does not."""
    p, sp, s0 = base
    s = s0.copy()
    s.jt_total = jt_total
    with gpu_case(p) as c:
        upload_all(c, s)
        c.sgs_stag()
        o.sgs_stag(sp, s)
        names = ["S11", "S13", "S33", "Nu_t", "txx", "txy", "tyy", "tzz"]
        if jt_total % 5 == 0:
            names += ["F_LM", "F_MM", "F_QN", "F_NN", "Cs_opt2", "Beta"]
        for n in names:
            assert_close(c, s, n, slice(1, p.nz), 1e-10)
        for n in ("txz", "tyz"):
            assert_close(c, s, n, slice(2, p.nz), 1e-10)


def test_interpolag(base):
    p, sp, s0 = base
    s = s0.copy()
    rng = np.random.default_rng(11)
    for n in ("F_LM", "F_MM", "F_QN", "F_NN"):
        s.f[n][:] = rng.uniform(0.5, 1.5, s.f[n].shape)
    with gpu_case(p) as c:
        upload_all(c, s)
        c.interpolag_Sdep()
        o.interpolag_Sdep(p, s)
        for n in ("F_LM", "F_NN"):
            assert_close(c, s, n, slice(1, p.nz + 1), 1e-13)


def test_divstress(base):
    p, sp, s0 = base
    s = s0.copy()
    with gpu_case(p) as c:
        upload_all(c, s)
        c.divstress_uv()
        c.divstress_w()
        o.divstress_uv(sp, s)
        o.divstress_w(sp, s)
        assert_close(c, s, "divtx", slice(1, p.nz), TOL, cols=p.ld)
        assert_close(c, s, "divty", slice(1, p.nz), TOL, cols=p.ld)
        assert_close(c, s, "divtz", slice(1, p.nz + 1), TOL)


def test_convec(base):
    p, sp, s0 = base
    s = s0.copy()
    with gpu_case(p) as c:
        upload_all(c, s)
        c.convec()
        o.convec(sp, s)
        assert_close(c, s, "RHSx", slice(1, p.nz), TOL)
        assert_close(c, s, "RHSy", slice(1, p.nz), TOL)
        assert_close(c, s, "RHSz", slice(1, p.nz + 1), TOL)


def test_press_stag_array_and_project(base):
    """Input u* is the projected field plus a non-solenoidal perturbation, as
    after the AB2 update (a divergence-free u* leaves only roundoff in the
    Poisson right-hand side)."""
    p, sp, s0 = base
    s = s0.copy()
    rng = np.random.default_rng(12)
    for n in ("u", "v", "w"):
        s.f[n][1:p.nz, :, :p.nx] += 0.5 * rng.standard_normal((p.nz - 1, p.ny, p.nx))
    with gpu_case(p) as c:
        upload_all(c, s)
        c.press_stag_array()
        o.press_stag_array(sp, s)
        assert_close(c, s, "p", slice(0, p.nz + 1), TOL)
        assert_close(c, s, "dpdx", slice(1, p.nz), TOL)
        assert_close(c, s, "dpdy", slice(1, p.nz), TOL)
        assert_close(c, s, "dpdz", slice(1, p.nz + 1), TOL)
        c.project()
        o.project(p, s)
        for n in ("u", "v", "w"):
            assert_close(c, s, n, slice(1, p.nz + 1), TOL)


def test_diagnostics(base):
    p, sp, s0 = base
    s = s0.copy()
    with gpu_case(p) as c:
        upload_all(c, s)
        assert abs(c.rmsdiv() - o.rmsdiv(p, s)) <= 1e-12 * max(o.rmsdiv(p, s), 1e-300) + 1e-300
        assert abs(c.energy() - o.energy(p, s)) <= 1e-12 * o.energy(p, s)
        assert abs(c.max_cfl() - o.max_cfl(p, s)) <= 1e-14
        assert abs(c.tau_wall_bot() - o.tau_wall_bot(p, s)) <= 1e-12 * o.tau_wall_bot(p, s)


def test_bad_arguments_fail_loudly():
    import lesgo_amd
    with pytest.raises(lesgo_amd.LesgoError):
        lesgo_amd.Case(lesgo_amd.Config(30, 32, 8))          # 30 = 2*3*5
    with gpu_case(o.Params(16, 16, 6)) as c:
        with pytest.raises(lesgo_amd.LesgoError):
            c.filt_da(99, 1, 2)
        with pytest.raises(ValueError):
            c.upload("u", np.zeros((3, 3, 3)))
