"""Instantaneous point and plane samples and the domain vorticity file.

Host-side writers for lesgo's instantaneous outputs (io.f90:534-678
``output_loop``, io.f90:681-1070 ``inst_write``), fed with fields downloaded
from the device.  Arrays are the library's [k][j][i] layout, k = 0..nz
(lbz = 0), i over ld columns; Fortran index (i, j, k) is [k][j - 1][i - 1].

- points (itype 1): ``output/vel.x-<x>.y-<y>.z-<z>.dat``, one list-directed
  record per output step appended: total_time and u, v, w_uv by trilinear
  interpolation (functions.f90:477-559) at the point;
- x-, y-, z-planes (itypes 3-5): ``output/vel.{x,y,z}-<loc>.<jt>.c0.bin``,
  three direct-access records (u, v, w_uv linearly interpolated to the plane,
  functions.f90:691-709);
- the domain vorticity (itype 2): ``output/vort.<jt>.c0.bin`` on the w grid,
  io.f90:823-858.

Cell indices and offsets follow output_init (io.f90:1284-1408) and cell_indx
(functions.f90:269-346): x and y locations are wrapped into [0, L) first.
z-planes and points are written when z(1) <= z < z(nz), the condition of the
reference's MPI build (one process, coord 0, the ``.c0`` names); its non-MPI
build writes no z-planes.  Where the reference reads the cell's upper
neighbour one past the last one (istart = Nx for an x-plane: the padding
column; jstart = Ny for a y-plane: past the array) the neighbour is taken
periodically (column 1 / row 1).  Locations in file names use the (f18.5)
edit of string_util.f90:61 (rounded half away from zero); point records are
written as Python reprs separated by spaces, one record per line, as
``check_ke.dat`` is.
"""
from __future__ import annotations

import math
import os
from decimal import ROUND_HALF_UP, Decimal

import numpy as np

THRESH = 1e-9      # cell_indx's boundary tolerance (functions.f90:295)


def fmt_real(r: float) -> str:
    """write(b, '(f18.5)') r, adjustl, trim (string_util.f90:61, 629-631)."""
    s = str(Decimal(float(r)).quantize(Decimal("0.00001"), rounding=ROUND_HALF_UP))
    if s.startswith("-") and Decimal(s) == 0:
        s = s[1:]
    return s


def wrap(px: float, L: float) -> float:
    """Fortran modulo(px, L) for L > 0: the result lies in [0, L)."""
    return px - math.floor(px / L) * L


def cell_indx(axis: str, d: float, px: float, n: int, L: float, z=None) -> int:
    """functions.f90:269-346 (Fortran, 1-based): 'i' / 'j' wrap px into [0, L)
    (pass it wrapped); 'k' uses z(1) and z(nz) (``z`` the level centres,
    indexed by k = 0..nz)."""
    if axis in ("i", "j"):
        if abs(px) / L < THRESH:
            return 1
        if abs(px - L) / L < THRESH:
            return n
        return math.floor(px / d) + 1
    nz = len(z) - 1
    if abs(px - z[nz]) / L < THRESH:
        return nz - 1
    return math.floor((px - z[1]) / d) + 1


def w_on_uv(w: np.ndarray, nz: int) -> np.ndarray:
    """interp_to_uv_grid (functions.f90:51-126) on levels 1..nz; level 0 is
    left as w's (the reference leaves it unset)."""
    r = w.copy()
    r[1:nz] = 0.5 * (w[2:nz + 1] + w[1:nz])
    r[nz] = r[nz - 1]
    return r


def on_w_grid(a: np.ndarray, nz: int) -> np.ndarray:
    """interp_to_w_grid (functions.f90:129-185), lbz = 0: levels 1..nz."""
    r = a.copy()
    r[1:nz + 1] = 0.5 * (a[0:nz] + a[1:nz + 1])
    return r


def lin(u1, u2, d, diff):
    """linear_interp_ss (functions.f90:691-709)."""
    return u1 + diff * (u2 - u1) / d


class Grid:
    """The uniform grid of grid.f90:44-128 (one process): nz levels above
    the wall (k = 0..nz stored), dx = lx / nx, dy = ly / ny, dz = lz / (nz - 1)
    unless given."""

    def __init__(self, nx, ny, nz, lx, ly, lz, dx=None, dy=None, dz=None):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.lx, self.ly, self.lz = lx, ly, lz
        self.dx = dx if dx is not None else lx / nx
        self.dy = dy if dy is not None else ly / ny
        self.dz = dz if dz is not None else lz / (nz - 1)
        self.z = np.array([(k - 0.5) * self.dz for k in range(nz + 1)])

    def x(self, i):            # Fortran i
        return (i - 1) * self.dx

    def y(self, j):
        return (j - 1) * self.dy


def trilinear(var: np.ndarray, g: Grid, xyz) -> float:
    """trilinear_interp (functions.f90:477-559) of var at a wrapped point."""
    i = cell_indx("i", g.dx, xyz[0], g.nx, g.lx)
    j = cell_indx("j", g.dy, xyz[1], g.ny, g.ly)
    k = cell_indx("k", g.dz, xyz[2], g.nz, g.lz, g.z)
    i1 = 1 if i + 1 == g.nx + 1 else i + 1
    j1 = 1 if j + 1 == g.ny + 1 else j + 1
    k1 = k + 1
    xd = xyz[0] - g.x(i)
    yd = xyz[1] - g.y(j)
    zd = xyz[2] - g.z[k]
    v = lambda ii, jj, kk: float(var[kk, jj - 1, ii - 1])
    u1 = v(i, j, k) + xd * (v(i1, j, k) - v(i, j, k)) / g.dx
    u2 = v(i, j1, k) + xd * (v(i1, j1, k) - v(i, j1, k)) / g.dx
    u3 = v(i, j, k1) + xd * (v(i1, j, k1) - v(i, j, k1)) / g.dx
    u4 = v(i, j1, k1) + xd * (v(i1, j1, k1) - v(i, j1, k1)) / g.dx
    u5 = u1 + yd * (u2 - u1) / g.dy
    u6 = u3 + yd * (u4 - u3) / g.dy
    return u5 + zd * (u6 - u5) / g.dz


def x_plane(var: np.ndarray, g: Grid, loc: float) -> np.ndarray:
    """[k - 1][j - 1] = var linearly interpolated to x = loc, k = 1..nz."""
    i = cell_indx("i", g.dx, loc, g.nx, g.lx)
    i1 = 1 if i == g.nx else i + 1
    return lin(var[1:g.nz + 1, :g.ny, i - 1], var[1:g.nz + 1, :g.ny, i1 - 1], g.dx, loc - g.x(i))


def y_plane(var: np.ndarray, g: Grid, loc: float) -> np.ndarray:
    """[k - 1][i - 1] = var linearly interpolated to y = loc, k = 1..nz."""
    j = cell_indx("j", g.dy, loc, g.ny, g.ly)
    j1 = 1 if j == g.ny else j + 1
    return lin(var[1:g.nz + 1, j - 1, :g.nx], var[1:g.nz + 1, j1 - 1, :g.nx], g.dy, loc - g.y(j))


def z_plane(var: np.ndarray, g: Grid, loc: float) -> np.ndarray:
    """[j - 1][i - 1] = var linearly interpolated to z = loc."""
    k = cell_indx("k", g.dz, loc, g.nz, g.lz, g.z)
    return lin(var[k, :g.ny, :g.nx], var[k + 1, :g.ny, :g.nx], g.dz, loc - g.z[k])


def in_z_range(g: Grid, zl: float) -> bool:
    return g.z[1] <= zl < g.z[g.nz]


def vorticity(d: dict, nz: int):
    """io.f90:823-835: vortx = dwdy - dvdz, vorty = dudz - dwdx, vortz =
    (dvdx - dudy) on the w grid with level 1 set to zero (coord 0)."""
    vx = d["dwdy"] - d["dvdz"]
    vy = d["dudz"] - d["dwdx"]
    vz = on_w_grid(d["dvdx"] - d["dudy"], nz)
    vz[1] = 0.0
    return vx, vy, vz


def _due(on, n0, n1, nskip, jt):
    return on and n0 <= jt <= n1 and (jt - n0) % nskip == 0


def _write_records(fname, arrays):
    with open(fname, "wb") as f:
        for a in arrays:
            f.write(np.ascontiguousarray(a).astype("<f8").tobytes())


class Sampler:
    """The point / plane schedule and locations of a ``LesgoConf``
    (output_loop, io.f90:585-676; output_init, io.f90:1323-1405)."""

    def __init__(self, c):
        self.c = c
        self.g = Grid(c.nx, c.ny, c.nz, c.L_x, c.L_y, c.L_z, c.dx, c.dy, c.dz)
        g = self.g
        self.xl = [wrap(x, g.lx) for x in c.xplane_loc] if c.xplane_calc else []
        self.yl = [wrap(y, g.ly) for y in c.yplane_loc] if c.yplane_calc else []
        self.zl = [z for z in c.zplane_loc if in_z_range(g, z)] if c.zplane_calc else []
        self.pts = ([(wrap(p[0], g.lx), wrap(p[1], g.ly), p[2]) for p in c.point_loc
                     if in_z_range(g, p[2])] if c.point_calc else [])

    def due(self, jt):
        c = self.c
        return {
            "point": _due(c.point_calc, c.point_nstart, c.point_nend, c.point_nskip, jt),
            "x": _due(c.xplane_calc, c.xplane_nstart, c.xplane_nend, c.xplane_nskip, jt),
            "y": _due(c.yplane_calc, c.yplane_nstart, c.yplane_nend, c.yplane_nskip, jt),
            "z": _due(c.zplane_calc, c.zplane_nstart, c.zplane_nend, c.zplane_nskip, jt),
        }

    def write(self, u, v, w, jt: int, total_time: float, out: str, due=None):
        """Writes what is due at step jt from the fields u, v, w."""
        due = due or self.due(jt)
        if not any(due.values()):
            return
        g = self.g
        wu = w_on_uv(w, g.nz)
        if due["point"]:
            for p in self.pts:
                fname = os.path.join(out, f"vel.x-{fmt_real(p[0])}.y-{fmt_real(p[1])}"
                                          f".z-{fmt_real(p[2])}.dat")
                vals = [total_time] + [trilinear(a, g, p) for a in (u, v, wu)]
                with open(fname, "a") as f:
                    f.write(" " + " ".join(repr(float(x)) for x in vals) + "\n")
        for key, locs, fn in (("x", self.xl, x_plane), ("y", self.yl, y_plane),
                              ("z", self.zl, z_plane)):
            if due[key]:
                for loc in locs:
                    fname = os.path.join(out, f"vel.{key}-{fmt_real(loc)}.{jt}.c0.bin")
                    _write_records(fname, [fn(a, g, loc) for a in (u, v, wu)])
