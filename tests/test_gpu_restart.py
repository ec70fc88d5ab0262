"""Restart through the reference's checkpoint files (vel.out, grid.out,
total_time.dat): a case written after 10 steps and read back continues
exactly (bit for bit) as the uninterrupted run, in Python and in the Fortran
driver, whose own final checkpoint reads back as the same state."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lesgo_amd", "lesgo_amd_main")
FIELDS = ("u", "v", "w", "RHSx", "F_LM", "Cs_opt2")


def test_restart_is_exact(tmp_path):
    import lesgo_amd
    from lesgo_amd.checkpoint import read_checkpoint, write_checkpoint
    cfg = lesgo_amd.Config(32, 32, 16)
    with lesgo_amd.Case(cfg) as a:
        a.seed_state(4)
        a.step(10)
        write_checkpoint(a, str(tmp_path))
        a.step(10)
        ref = {n: a.download(n) for n in FIELDS + ("p",)}
    with lesgo_amd.Case(cfg) as b:
        jt, t = read_checkpoint(b, str(tmp_path))
        assert jt == 10 and abs(t - 10 * cfg.dt) < 1e-15
        b.step(10)
        for n in FIELDS + ("p",):
            d = np.max(np.abs(b.download(n)[1:17] - ref[n][1:17]))
            assert d == 0.0, (n, d)
    # the Fortran driver restarts from the same files and checkpoints again
    with open(tmp_path / "lesgo_amd.nml", "w") as f:
        f.write(f"&lesgo Nx=32, Ny=32, Nz=16, nsteps=10, wbase=10, domain_nskip=10, "
                f"fused=.true., checkpoint_data=.true., path='{tmp_path}/' /\n")
    os.makedirs(tmp_path / "output", exist_ok=True)
    r = subprocess.run([EXE, str(tmp_path / "lesgo_amd.nml")], cwd=tmp_path,
                       capture_output=True, text=True, timeout=180)
    print(r.stdout[-800:], r.stderr[-800:], flush=True)
    assert r.returncode == 0 and "Simulation complete" in r.stdout
    n = 32 * 32 * 17
    vel = np.fromfile(tmp_path / "output" / "vel.20.c0.bin", dtype=np.float64)
    np.testing.assert_array_equal(vel[:n].reshape(17, 32, 32), ref["u"][1:18, :, :32])
    with lesgo_amd.Case(cfg) as c:
        jt, _ = read_checkpoint(c, str(tmp_path))
        assert jt == 20
        for name in FIELDS:
            d = np.max(np.abs(c.download(name)[1:17] - ref[name][1:17]))
            assert d == 0.0, (name, d)
