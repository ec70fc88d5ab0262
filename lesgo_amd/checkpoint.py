"""The reference's restart files for a device-resident case.

``write_checkpoint`` writes what ``checkpoint`` writes (io.f90:1173-1267) for
one process: ``vel.out`` (one unformatted sequential record holding u, v, w,
RHSx, RHSy, RHSz, Cs_opt2, F_LM, F_MM, F_QN, F_NN, each ``(ld, ny, 1:nz)``
float64), ``grid.out`` (nproc, Nx, Ny, Nz, L_x, L_y, L_z) and
``total_time.dat`` (jt_total, total_time, total_time_dim, dt, cfl).
``read_checkpoint`` is ``ic_file`` (initial.f90:226-239) plus the
``total_time.dat`` counters and clock (io.f90:62-95): with use_cfl_dt the
file's dt is the previous step's, and a new run (jt_total = 0) or a cfl that
changed by more than 1 % restarts with an Euler step (initialize.f90:191-200).  Unformatted records use 4-byte little-endian
length markers; a record longer than 2**31 - 9 bytes is split into
subrecords whose markers are negative except the last leading and the
first trailing one (gfortran's convention).
"""
from __future__ import annotations

import os
import struct

import numpy as np

from .case import Case

CHECKPOINT_FIELDS = ("u", "v", "w", "RHSx", "RHSy", "RHSz", "Cs_opt2", "F_LM", "F_MM",
                     "F_QN", "F_NN")
_MAXSUB = 2 ** 31 - 9


def write_record(f, payload: bytes):
    """One Fortran unformatted sequential record."""
    n = len(payload)
    if n <= _MAXSUB:
        f.write(struct.pack("<i", n))
        f.write(payload)
        f.write(struct.pack("<i", n))
        return
    pos, first = 0, True
    while pos < n:
        chunk = payload[pos:pos + _MAXSUB]
        last = pos + len(chunk) >= n
        f.write(struct.pack("<i", len(chunk) if last else -len(chunk)))
        f.write(chunk)
        f.write(struct.pack("<i", len(chunk) if first else -len(chunk)))
        pos += len(chunk)
        first = False


def read_record(f) -> bytes:
    out = []
    while True:
        (lead,) = struct.unpack("<i", f.read(4))
        out.append(f.read(abs(lead)))
        f.read(4)
        if lead >= 0:
            return b"".join(out)


def write_checkpoint(case: Case, path: str = ".", member: int = 0, fname: str = "vel.out",
                     total_time_dim: float | None = None):
    """vel.out, grid.out and total_time.dat of member `member` into `path`
    (total_time_dim: the dimensional time column, total_time when not
    given)."""
    cfg = case.cfg
    nz = case.nz
    # the files are the state from here on, in this run as in a restart
    case.resync()
    jt, jt_total = case.counters()
    parts = [np.ascontiguousarray(case.download(n, member)[1:nz + 1]) for n in CHECKPOINT_FIELDS]
    with open(os.path.join(path, fname), "wb") as f:
        write_record(f, b"".join(a.astype("<f8").tobytes() for a in parts))
    # io.f90:1220 writes the reference's Nz, which is param's nz (Fortran
    # names ignore case): nz_tot/nproc + 1 = the input level count + 1 on one
    # process (input_util.f90:196-201), the value initial.f90:211-214 checks
    with open(os.path.join(path, "grid.out"), "wb") as f:
        write_record(f, struct.pack("<4i3d", 1, cfg.nx, cfg.ny, nz, cfg.lx, cfg.ly, cfg.lz))
    # the cfl written is the target with use_cfl_dt, the measured one
    # otherwise (io.f90:1232-1238)
    clk = case.clock(member)
    cfl = cfg.cfl if cfg.use_cfl_dt else case.max_cfl(member)
    t, dt = clk["total_time"], clk["dt"]
    td = t if total_time_dim is None else total_time_dim
    with open(os.path.join(path, "total_time.dat"), "w") as f:
        f.write(f" {jt_total:d} {t!r} {td!r} {dt!r} {cfl!r}\n")


def read_checkpoint(case: Case, path: str = ".", member: int = 0, fname: str = "vel.out",
                    set_counters: bool = True):
    """Load a checkpoint into member `member`; returns (jt_total, total_time)
    from total_time.dat (0, 0.0 if absent) and sets the case's counters.
    A restart runs the file-start schedule (the reference sets initu =
    .true. when the file exists), so the case must be made with initu."""
    if not case.cfg.initu:
        raise ValueError("a restarted case runs the file-start schedule: make it with initu=True")
    nz = case.nz
    shp = (nz, case.ny, case.ld)
    with open(os.path.join(path, fname), "rb") as f:
        raw = read_record(f)
    n = int(np.prod(shp))
    if len(raw) != 8 * n * len(CHECKPOINT_FIELDS):
        raise ValueError(f"{fname}: {len(raw)} bytes, expected "
                         f"{8 * n * len(CHECKPOINT_FIELDS)} for this grid")
    allv = np.frombuffer(raw, dtype="<f8")
    for q, name in enumerate(CHECKPOINT_FIELDS):
        a = case.download(name, member)
        a[1:nz + 1] = allv[q * n:(q + 1) * n].reshape(shp)
        if name in ("u", "v", "w"):
            a[0] = -1234567890.0
        case.upload(name, a, member)
    cfg = case.cfg
    jt_total, t, dt_r, cfl_r = 0, 0.0, cfg.dt, cfg.cfl
    tt = os.path.join(path, "total_time.dat")
    if os.path.exists(tt):
        w = open(tt).read().split()
        jt_total, t, dt_r, cfl_r = int(w[0]), float(w[1]), float(w[3]), float(w[4])
    fresh = cfg.use_cfl_dt and (jt_total == 0 or abs((cfl_r - cfg.cfl) / cfg.cfl) > 1e-2)
    case.set_clock(dt_r if cfg.use_cfl_dt else cfg.dt, t, fresh, member)
    if set_counters:
        case.set_counters(jt_total, jt_total)
    return jt_total, t


# tavg.out: the averaging time, then the running sums in this order (p is
# summed but not written, time_average.f90:665-716), each (nx, ny, lbz:nz)
TAVG_CHECKPOINT = ("u", "v", "w", "u_w", "v_w", "w_uv", "u2", "v2", "w2", "uv", "uw", "vw",
                   "txx", "tyy", "tzz", "txy", "txz", "tyz", "fx", "fy", "fz", "cs_opt2",
                   "vortx", "vorty", "vortz")


def write_tavg_checkpoint(case: Case, path: str = ".", member: int = 0, fname: str = "tavg.out"):
    """tavg%checkpoint (time_average.f90:665-716): one record with the
    averaging time, then one per array."""
    _, initialized = case.stats_info()
    if not initialized:
        return False
    total_time = case.stats_time(member)
    with open(os.path.join(path, fname), "wb") as f:
        write_record(f, struct.pack("<d", total_time))
        for n in TAVG_CHECKPOINT:
            write_record(f, np.ascontiguousarray(case.stats_sums(n, member)).astype("<f8").tobytes())
    return True


def read_tavg_checkpoint(case: Case, path: str = ".", member: int = 0, fname: str = "tavg.out"):
    """tavg%init with an existing tavg.out (time_average.f90:113-166): the
    sums continue from the file (the pressure sum from zero, as the file has
    none); call after ``stats_enable``.  Returns False without a file."""
    fn = os.path.join(path, fname)
    if not os.path.exists(fn):
        return False
    shp = (case.nz + 1, case.ny, case.cfg.nx)
    with open(fn, "rb") as f:
        (total_time,) = struct.unpack("<d", read_record(f))
        sums = {n: np.frombuffer(read_record(f), dtype="<f8").reshape(shp) for n in TAVG_CHECKPOINT}
    case.stats_restore(total_time, sums, member)
    return True
