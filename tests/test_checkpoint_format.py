"""The restart file format: Fortran unformatted sequential records written
and read by lesgo_amd.checkpoint, checked against the Fortran compiler's own
I/O (a small program built with amdflang reads the Python-written record and
writes one Python reads back), and the subrecord layout of long records."""
import io
import os
import shutil
import subprocess

import numpy as np
import pytest

from lesgo_amd import checkpoint as ck

FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")

PROG = """
program rw
implicit none
integer, parameter :: ld = 10, ny = 6, nz = 5
real(8) :: a(ld, ny, 0:nz), b(ld, ny, 0:nz)
open(12, file='in.rec', form='unformatted', status='old')
read(12) a(:, :, 1:nz), b(:, :, 1:nz)
close(12)
write(*, '(2es25.16)') sum(a(:, :, 1:nz)), sum(b(:, :, 1:nz) * 2)
open(13, file='out.rec', form='unformatted', status='replace')
write(13) b(:, :, 1:nz) * 3, a(:, :, 1:nz)
close(13)
end program rw
"""


def test_subrecords_round_trip(monkeypatch):
    monkeypatch.setattr(ck, "_MAXSUB", 100)
    payload = bytes(range(256)) + bytes(94)       # 350 bytes -> 100, 100, 100, 50
    f = io.BytesIO()
    ck.write_record(f, payload)
    raw = f.getvalue()
    marks = []
    pos = 0
    while pos < len(raw):
        lead = int.from_bytes(raw[pos:pos + 4], "little", signed=True)
        n = abs(lead)
        trail = int.from_bytes(raw[pos + 4 + n:pos + 8 + n], "little", signed=True)
        marks.append((lead, trail))
        pos += n + 8
    assert marks == [(-100, 100), (-100, -100), (-100, -100), (50, -50)]
    f.seek(0)
    assert ck.read_record(f) == payload


@pytest.mark.skipif(not os.path.exists(FC), reason="no Fortran compiler")
def test_records_interoperate_with_fortran_io(tmp_path):
    src = tmp_path / "rw.f90"
    src.write_text(PROG)
    exe = tmp_path / "rw"
    subprocess.run([FC, "-O1", str(src), "-o", str(exe)], check=True, cwd=tmp_path,
                   capture_output=True, timeout=120)
    rng = np.random.default_rng(0)
    ld, ny, nz = 10, 6, 5
    a = rng.standard_normal((nz, ny, ld))          # Fortran a(:, :, 1:nz)
    b = rng.standard_normal((nz, ny, ld))
    with open(tmp_path / "in.rec", "wb") as f:
        ck.write_record(f, a.astype("<f8").tobytes() + b.astype("<f8").tobytes())
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    sa, sb = (float(x) for x in r.stdout.split())
    np.testing.assert_allclose([sa, sb], [a.sum(), 2 * b.sum()], rtol=1e-13)
    with open(tmp_path / "out.rec", "rb") as f:
        raw = np.frombuffer(ck.read_record(f), dtype="<f8")
    n = a.size
    np.testing.assert_array_equal(raw[:n].reshape(a.shape), 3 * b)
    np.testing.assert_array_equal(raw[n:].reshape(a.shape), a)
    shutil.rmtree(tmp_path, ignore_errors=True)


class _StatsStub:
    """The part of a Case the tavg.out writer and reader use."""

    class _Cfg:
        nx = 6

    def __init__(self):
        self.nz, self.ny, self.cfg = 4, 5, self._Cfg()
        self.restored = None

    def stats_info(self):
        return 0.375, True

    def stats_time(self, member=0):
        return 0.375 + member

    def stats_sums(self, name, member=0):
        q = ck.TAVG_CHECKPOINT.index(name) if name in ck.TAVG_CHECKPOINT else 99
        return np.arange(5 * 5 * 6, dtype=float).reshape(5, 5, 6) + 1000.0 * q

    def stats_restore(self, total_time, sums, member=0):
        self.restored = (total_time, sums)


def test_tavg_checkpoint_layout_and_round_trip(tmp_path):
    """tavg.out (time_average.f90:665-716): a record with the averaging time,
    then one per running sum in the reference's order with no pressure
    record, each (nx, ny, 0:nz) doubles; reading it back restores exactly
    those sums (the pressure sum absent, so it restarts from zero)."""
    c = _StatsStub()
    assert ck.write_tavg_checkpoint(c, str(tmp_path))
    size = os.path.getsize(tmp_path / "tavg.out")
    recs = []
    with open(tmp_path / "tavg.out", "rb") as f:
        while f.tell() < size:
            recs.append(ck.read_record(f))
    assert len(recs) == 26
    assert np.frombuffer(recs[0], "<f8").tolist() == [0.375]
    assert all(len(r) == 5 * 5 * 6 * 8 for r in recs[1:])
    assert len(ck.TAVG_CHECKPOINT) == 25 and "p" not in ck.TAVG_CHECKPOINT
    assert ck.TAVG_CHECKPOINT[:6] == ("u", "v", "w", "u_w", "v_w", "w_uv")
    np.testing.assert_array_equal(np.frombuffer(recs[4], "<f8").reshape(5, 5, 6),
                                  c.stats_sums("u_w"))
    assert ck.read_tavg_checkpoint(c, str(tmp_path))
    t, sums = c.restored
    assert t == 0.375 and set(sums) == set(ck.TAVG_CHECKPOINT)
    for n in ck.TAVG_CHECKPOINT:
        np.testing.assert_array_equal(sums[n], c.stats_sums(n))
    assert not ck.read_tavg_checkpoint(c, str(tmp_path / "none"))


GRID_PROG = """
program grid
implicit none
! the reference's param declares Nz only; nz in the code is the same name
! (Fortran ignores case), set to nz_tot/nproc + 1 (input_util.f90:196-201)
integer :: Nx = 8, Ny = 6, Nz = 4
real(8) :: L_x = 6.283185307179586d0, L_y = 3.141592653589793d0, L_z = 1d0
nz = Nz + 1
open(11, file='grid.out', form='unformatted', status='replace')
write(11) 1, Nx, Ny, Nz, L_x, L_y, L_z
close(11)
end program grid
"""


class _CaseStub:
    """The part of a Case the vel.out / grid.out / total_time.dat writer uses."""

    def __init__(self, nx, ny, Nz):
        from lesgo_amd.case import Config
        self.cfg = Config(nx, ny, Nz)
        self.nz, self.ny, self.ld = Nz + 1, ny, nx + 2

    def resync(self):
        pass

    def counters(self):
        return 7, 7

    def download(self, name, member=0):
        return np.zeros((self.nz + 1, self.ny, self.ld))

    def clock(self, member=0):
        return {"total_time": 0.25, "dt": 2e-4}

    def max_cfl(self, member=0):
        return 0.1


@pytest.mark.skipif(not os.path.exists(FC), reason="no Fortran compiler")
def test_grid_out_holds_the_references_nz(tmp_path):
    """grid.out (io.f90:1220) holds (nproc, Nx, Ny, Nz, L_x, L_y, L_z) with
    the reference's Nz, i.e. param's nz = the input level count + 1 on one
    process; the bytes equal those of the Fortran statement compiled with
    the same case-insensitive names."""
    src = tmp_path / "grid.f90"
    src.write_text(GRID_PROG)
    d = tmp_path / "f"
    d.mkdir()
    subprocess.run([FC, "-O1", str(src), "-o", str(d / "grid")], check=True, cwd=d,
                   capture_output=True, timeout=120)
    subprocess.run([str(d / "grid")], check=True, cwd=d, capture_output=True, timeout=60)
    py = tmp_path / "py"
    py.mkdir()
    ck.write_checkpoint(_CaseStub(8, 6, 4), str(py))
    with open(py / "grid.out", "rb") as f:
        rec = ck.read_record(f)
    import struct
    assert struct.unpack("<4i3d", rec)[:4] == (1, 8, 6, 5)
    assert (py / "grid.out").read_bytes() == (d / "grid.out").read_bytes()
