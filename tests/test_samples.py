"""Point and plane samples and the domain vorticity file (lesgo_amd.samples,
io.f90:534-1070) on the host: interpolation pinned by fields linear in each
coordinate (trilinear and linear interpolation are exact on them, so a wrong
cell index, offset, spacing or level fails), the reference's cell_indx edge
cases, the (f18.5) file-name edit, the record layouts and the schedule."""
import os

import numpy as np
import pytest

from lesgo_amd import samples as sm
from lesgo_amd.conf import LesgoConf

NX, NY, NZ = 16, 12, 9          # nz levels above the wall (k = 0..NZ stored)
LX, LY, LZ = 2.0, 1.2, 1.0          # dx, dy, dz all differ


def grid():
    return sm.Grid(NX, NY, NZ, LX, LY, LZ)


def f_lin(x, y, z):
    """linear in each coordinate separately"""
    return 0.3 + 1.7 * x - 0.9 * y + 2.3 * z + 0.4 * x * y - 0.6 * y * z + 0.25 * x * z + 0.1 * x * y * z


def field(g, fn, zshift=0.0):
    """fn on the uv grid (x_i, y_j, z_k - zshift), [k][j][i] with ld = nx + 2"""
    a = np.full((g.nz + 1, g.ny, g.nx + 2), 1e30)
    z = g.z - zshift
    x = np.arange(g.nx) * g.dx
    y = np.arange(g.ny) * g.dy
    a[:, :, :g.nx] = fn(x[None, None, :], y[None, :, None], z[:, None, None])
    return a


def test_fmt_real_f18_5():
    assert sm.fmt_real(0.5) == "0.50000"
    assert sm.fmt_real(3.141592653589793) == "3.14159"
    assert sm.fmt_real(100.0) == "100.00000"
    assert sm.fmt_real(0.0) == "0.00000"
    assert sm.fmt_real(1.0 / 64) == "0.01563"       # an exact tie rounds away from zero
    assert sm.fmt_real(-1.0 / 64) == "-0.01563"
    assert sm.fmt_real(2.999996) == "3.00000"


def test_cell_indx_edges():
    g = grid()
    assert sm.cell_indx("i", g.dx, 0.0, NX, LX) == 1
    assert sm.cell_indx("i", g.dx, LX * (1 - 1e-12), NX, LX) == NX     # within 1e-9 of L
    assert sm.cell_indx("i", g.dx, 1.5 * g.dx, NX, LX) == 2
    assert sm.cell_indx("j", g.dy, sm.wrap(-0.5 * g.dy, LY), NY, LY) == NY
    assert sm.wrap(-0.5 * g.dy, LY) == pytest.approx(LY - 0.5 * g.dy)
    assert sm.wrap(LX + 0.25, LX) == pytest.approx(0.25)
    # k: z(1) = dz/2 is cell 1; z(nz) is cell nz - 1
    assert sm.cell_indx("k", g.dz, g.z[1], NZ, LZ, g.z) == 1
    assert sm.cell_indx("k", g.dz, g.z[NZ], NZ, LZ, g.z) == NZ - 1
    assert sm.cell_indx("k", g.dz, g.z[3] + 0.2 * g.dz, NZ, LZ, g.z) == 3


def test_trilinear_exact_on_multilinear_fields():
    g = grid()
    u = field(g, f_lin)
    rng = np.random.default_rng(3)
    for _ in range(50):
        x = rng.uniform(0, LX - g.dx)            # cells that do not wrap
        y = rng.uniform(0, LY - g.dy)
        z = rng.uniform(g.z[1], g.z[NZ])
        assert sm.trilinear(u, g, (x, y, z)) == pytest.approx(f_lin(x, y, z), rel=1e-13, abs=1e-13)
    # the last cell takes its upper neighbour periodically (column / row 1)
    c = field(g, lambda x, y, z: 1.0 + 0 * x + 2.0 * z + 0 * y)
    assert sm.trilinear(c, g, (LX - 0.3 * g.dx, LY - 0.2 * g.dy, 0.5)) == pytest.approx(2.0)


def test_planes_exact_and_layout():
    g = grid()
    u = field(g, f_lin)
    yj = np.arange(NY) * g.dy
    xi = np.arange(NX) * g.dx
    zk = g.z[1:NZ + 1]
    xl = 0.37 * LX
    pl = sm.x_plane(u, g, xl)
    assert pl.shape == (NZ, NY)                  # record: j fastest, then k
    np.testing.assert_allclose(pl, f_lin(xl, yj[None, :], zk[:, None]), rtol=1e-13, atol=1e-13)
    yl = 0.61 * LY
    pl = sm.y_plane(u, g, yl)
    assert pl.shape == (NZ, NX)
    np.testing.assert_allclose(pl, f_lin(xi[None, :], yl, zk[:, None]), rtol=1e-13, atol=1e-13)
    zl = 0.47
    pl = sm.z_plane(u, g, zl)
    assert pl.shape == (NY, NX)
    np.testing.assert_allclose(pl, f_lin(xi[None, :], yj[:, None], zl), rtol=1e-13, atol=1e-13)


def test_w_interpolations():
    g = grid()
    # w on the w grid (zw_k = z_k - dz/2): w_uv(k) is w at z_k for k < nz, w_uv(nz) = w_uv(nz - 1)
    w = field(g, f_lin, zshift=0.5 * g.dz)
    wu = sm.w_on_uv(w, NZ)
    ref = field(g, f_lin)
    np.testing.assert_allclose(wu[1:NZ, :, :NX], ref[1:NZ, :, :NX], rtol=1e-13, atol=1e-13)
    np.testing.assert_array_equal(wu[NZ], wu[NZ - 1])
    # uv-grid quantity to the w grid: level k is the mean of k - 1 and k
    a = field(g, f_lin)
    aw = sm.on_w_grid(a, NZ)
    np.testing.assert_allclose(aw[1:, :, :NX], field(g, f_lin, zshift=0.5 * g.dz)[1:, :, :NX],
                               rtol=1e-13, atol=1e-13)


def test_vorticity_formula():
    g = grid()
    rng = np.random.default_rng(5)
    d = {n: rng.standard_normal((NZ + 1, NY, NX)) for n in
         ("dudy", "dudz", "dvdx", "dvdz", "dwdx", "dwdy")}
    vx, vy, vz = sm.vorticity(d, NZ)
    np.testing.assert_array_equal(vx, d["dwdy"] - d["dvdz"])
    np.testing.assert_array_equal(vy, d["dudz"] - d["dwdx"])
    wz = d["dvdx"] - d["dudy"]
    np.testing.assert_array_equal(vz[1], 0.0)
    np.testing.assert_allclose(vz[2:], 0.5 * (wz[1:NZ] + wz[2:]), rtol=0, atol=0)


def _conf(tmp_path):
    c = LesgoConf()
    c.nx, c.ny, c.nz = NX, NY, NZ
    c.L_x, c.L_y, c.L_z = LX, LY, LZ
    c.dx, c.dy, c.dz = LX / NX, LY / NY, LZ / (NZ - 1)
    c.point_calc, c.point_nstart, c.point_nend, c.point_nskip = True, 2, 10, 4
    c.point_loc = [(0.25, 0.5, 0.3), (LX + 0.1, -0.2, 0.61), (0.5, 0.5, 5.0)]   # the last: above z(nz)
    c.xplane_calc, c.xplane_nstart, c.xplane_nend, c.xplane_nskip = True, 3, 3, 1
    c.xplane_loc = [0.7]
    c.yplane_calc, c.yplane_nstart, c.yplane_nend, c.yplane_nskip = True, 3, 3, 1
    c.yplane_loc = [1.1]
    c.zplane_calc, c.zplane_nstart, c.zplane_nend, c.zplane_nskip = True, 3, 3, 1
    c.zplane_loc = [0.3, 0.0]                  # 0.0 is below z(1): not written
    return c


def test_sampler_files(tmp_path):
    c = _conf(tmp_path)
    s = sm.Sampler(c)
    g = s.g
    u = field(g, f_lin)
    v = field(g, lambda x, y, z: 2.0 * f_lin(x, y, z))
    w = field(g, lambda x, y, z: -f_lin(x, y, z), zshift=0.5 * g.dz)
    assert [k for k, on in s.due(2).items() if on] == ["point"]
    assert [k for k, on in s.due(3).items() if on] == ["x", "y", "z"]
    assert not any(s.due(4).values()) and s.due(6)["point"] and not s.due(14)["point"]
    for jt, t in ((2, 0.125), (3, 0.25), (6, 0.5)):
        s.write(u, v, w, jt, t, str(tmp_path))
    names = sorted(os.listdir(tmp_path))
    assert names == sorted(["vel.x-0.25000.y-0.50000.z-0.30000.dat",
                            "vel.x-0.10000.y-1.00000.z-0.61000.dat",
                            "vel.x-0.70000.3.c0.bin", "vel.y-1.10000.3.c0.bin",
                            "vel.z-0.30000.3.c0.bin"])
    rec = np.loadtxt(tmp_path / "vel.x-0.25000.y-0.50000.z-0.30000.dat")
    assert rec.shape == (2, 4)
    np.testing.assert_array_equal(rec[:, 0], [0.125, 0.5])
    fv = f_lin(0.25, 0.5, 0.3)
    np.testing.assert_allclose(rec[0, 1:], [fv, 2 * fv, -fv], rtol=1e-12)
    raw = np.fromfile(tmp_path / "vel.x-0.70000.3.c0.bin", dtype="<f8")
    assert raw.size == 3 * NY * NZ
    np.testing.assert_allclose(raw[NY * NZ:2 * NY * NZ].reshape(NZ, NY),
                               2 * sm.x_plane(u, g, 0.7), rtol=1e-14)
    raw = np.fromfile(tmp_path / "vel.z-0.30000.3.c0.bin", dtype="<f8")
    assert raw.size == 3 * NX * NY
    wu = sm.w_on_uv(w, NZ)
    np.testing.assert_allclose(raw[2 * NX * NY:].reshape(NY, NX), sm.z_plane(wu, g, 0.3), rtol=1e-14)
