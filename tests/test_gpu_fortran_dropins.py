"""The Fortran drop-ins of the reference procedures the time loop does not
call by name (lesgo_amd/fortran/dropins.f90): derivatives::ddx, ddy, ddxy,
test_filtermodule::test_filter and test_test_filter, the lagrange_Sdep and
interpolag_Sdep externals, tridag_array(a, b, c, r, u) and
turbines::turbines_forcing, each called through its Fortran interface on
the state the program wrote, against the oracle applied to that same state."""
import os
import subprocess

import numpy as np
import pytest

from oracle import lesgo_oracle as o
from gpu_util import rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lesgo_amd", "lesgo_amd_dropins")
NX, NY, NZ = 32, 24, 12
FARM = [o.TurbineSpec(2.0, 1.0, 0.3, 0.3, 0.0, 0.0, 0.0, 1.33),
        o.TurbineSpec(4.5, 2.2, 0.3, 0.3, 0.0, 0.0, 0.0, 1.33)]


class _Reader:
    def __init__(self, path):
        self.a = np.fromfile(path, dtype="<f8")
        self.pos = 0

    def take(self, *shape):
        n = int(np.prod(shape))
        out = self.a[self.pos:self.pos + n].reshape(shape)
        self.pos += n
        return out

    def done(self):
        return self.pos == self.a.size


def test_fortran_dropins_match_oracle(tmp_path):
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=180)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "dropins complete" in r.stdout
    p = o.Params(NX, NY, NZ)
    sp = o.Spectral(p)
    nz, ld, lh = p.nz, p.ld, p.lh
    F = (nz + 1, NY, ld)
    st = _Reader(tmp_path / "state.bin")
    s = o.State.zeros(p)
    for n in o.FIELDS:
        s.f[n][:] = st.take(*F)
    assert st.done()
    s.jt = s.jt_total = 3
    out = _Reader(tmp_path / "out.bin")
    X = np.s_[:, :, :p.nx]
    K = np.s_[1:nz + 1]

    def close(name, got, want, tol):
        e = rel_l2(got, want)
        print(f"  {name}: {e:.2e}", flush=True)
        assert e <= tol, (name, e)

    dudx, dvdy, dwdx, dwdy = (out.take(*F) for _ in range(4))
    close("ddx", dudx[K][X], o.ddx(sp, s.u)[K][X], 1e-12)
    close("ddy", dvdy[K][X], o.ddy(sp, s.v)[K][X], 1e-12)
    ex, ey = o.ddxy(sp, s.w)
    close("ddxy x", dwdx[K][X], ex[K][X], 1e-12)
    close("ddxy y", dwdy[K][X], ey[K][X], 1e-12)
    tf = out.take(NY, ld)
    close("test_filter", tf[:, :p.nx], sp.test_filter(s.u[2])[:, :p.nx], 1e-12)
    ttf = out.take(NY, ld)
    close("test_test_filter", ttf[:, :p.nx], sp.test_test_filter(s.v[3])[:, :p.nx], 1e-12)
    tu = out.take(nz + 1, NY, ld)
    ta, tb, tc = (out.take(nz + 1, NY, lh) for _ in range(3))
    tr = out.take(nz + 1, NY, ld)
    uc = o.tridag_array(p, ta, tb, tc, tr[..., 0::2] + 1j * tr[..., 1::2])
    m = np.zeros((nz + 1, NY, lh - 1), dtype=bool)
    m[0] = True
    m[1:] = True
    m[1:, NY // 2] = False
    m[1:, 0, 0] = False
    close("tridag_array re", tu[..., 0:2 * (lh - 1):2][m], uc.real[..., :lh - 1][m], 1e-13)
    close("tridag_array im", tu[..., 1:2 * (lh - 1):2][m], uc.imag[..., :lh - 1][m], 1e-13)
    si = s.copy()
    o.interpolag_Sdep(p, si)
    for n in ("F_LM", "F_MM", "F_QN", "F_NN"):
        close(f"interpolag {n}", out.take(*F)[K][X], si.f[n][K][X], 1e-13)
    sl = s.copy()
    sl.jt = sl.jt_total = p.cs_count
    o.lagrange_Sdep(sp, sl)
    for n in ("F_LM", "F_MM", "F_QN", "F_NN", "Cs_opt2", "Beta"):
        close(f"lagrange {n}", out.take(*F)[K][X], sl.f[n][K][X], 1e-10)
    stb = s.copy()
    tb_ = o.Turbines(p, FARM)
    tb_.forcing(stb)
    for n in ("fxa", "fya", "fza"):
        got = out.take(*F)
        if n == "fya":
            assert np.abs(got[K][X]).max() == 0.0 and np.abs(stb.f[n][K][X]).max() == 0.0
        else:
            close(f"turbines_forcing {n}", got[K][X], stb.f[n][K][X], 1e-10)
    assert out.done()
