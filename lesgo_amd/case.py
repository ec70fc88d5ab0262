"""Host-side mirror of the reference's time-step interface.

``Case`` owns one device-resident LES case (the sim_param / sgs_param
module state of the reference, sim_param.f90:31-93, sgs_param.f90:29-184)
and exposes one method per replaced subroutine, under the reference's names:
``filt_da``, ``ddz_uv``, ``ddz_w``, ``wallstress``, ``sgs_stag``,
``lagrange_Sdep``, ``interpolag_Sdep``, ``divstress_uv``, ``divstress_w``,
``convec``, ``press_stag_array``, ``project``, ``rmsdiv``, ``energy`` and
``turbines_forcing``; ``step`` is one pass of the time loop
(main.f90:146-344).  Host arrays use the reference layout: Fortran
``f(ld, ny, 0:nz)`` == numpy ``[nz+1, ny, ld]`` float64.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import FIELD_ID, FIELDS, LG_FP64, LG_MIXED, LgParams, check, lib

PI = 3.1415926535897932384626433


@dataclass
class Config:
    """Run parameters (param.f90 subset used on the time step)."""
    nx: int
    ny: int
    Nz: int
    lx: float = 2.0 * PI
    ly: float = PI
    lz: float = 1.0
    dt: float = 2.0e-4
    zo: float = 1.0e-4
    vonk: float = 0.4
    cs_count: int = 5
    use_mean_p_force: bool = True
    mean_p_force_x: float = 1.0
    mean_p_force_y: float = 0.0
    nu: float = 0.0
    precision: str = "fp64"          # "fp64" | "mixed"
    # initu: True = state from a file (dynamic model from the start); False =
    # fresh start (Euler first step, Cs_opt2 = 0.03 until dyn_init steps)
    initu: bool = True
    dyn_init: int = 100
    # use_cfl_dt: dt = cfl / max(|u|/dx, |v|/dy, |w|/dz) each step, on the
    # device (cfl_util.f90:73-111); dt above is then only the starting value
    use_cfl_dt: bool = False
    cfl: float = 0.05

    def to_c(self) -> LgParams:
        return LgParams(self.nx, self.ny, self.Nz, self.lx, self.ly, self.lz,
                        self.dt, self.zo, self.vonk, self.cs_count,
                        int(self.use_mean_p_force), self.mean_p_force_x,
                        self.mean_p_force_y, self.nu,
                        LG_FP64 if self.precision == "fp64" else LG_MIXED,
                        int(self.initu), self.dyn_init, int(self.use_cfl_dt),
                        self.cfl)


def _fid(f) -> int:
    return f if isinstance(f, int) else FIELD_ID[f]


class Case:
    """Device-resident LES case(s) behind the lesgo_amd C ABI.

    ``members > 1`` makes a batch of independent cases of one grid (each
    with its own fields, z0, mean pressure gradient and disks) that every
    step advances together; per-member methods take ``member``."""

    def __init__(self, cfg: Config, members: int = 1):
        self.cfg = cfg
        h = C.c_void_p()
        check(lib().lg_create_batch(C.byref(cfg.to_c()), members, C.byref(h)), "lg_create_batch")
        self._h = h
        nz, ld, ny = C.c_int(), C.c_int(), C.c_int()
        check(lib().lg_sizes(h, C.byref(nz), C.byref(ld), C.byref(ny)), "lg_sizes")
        self.nz, self.ld, self.ny = nz.value, ld.value, ny.value
        self.shape = (self.nz + 1, self.ny, self.ld)
        self.members = members
        self.nturb = [0] * members

    def close(self):
        if getattr(self, "_h", None):
            lib().lg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- module state ----------------------------------------------------
    def upload(self, name, arr, member: int = 0):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"{name}: expected shape {self.shape}, got {a.shape}")
        check(lib().lg_upload_member(self._h, member, _fid(name), a.ctypes.data_as(_lib._D)),
              f"upload {name}")

    def download(self, name, member: int = 0):
        a = np.empty(self.shape)
        check(lib().lg_download_member(self._h, member, _fid(name), a.ctypes.data_as(_lib._D)),
              f"download {name}")
        return a

    def seed_state(self, seed: int, noise: float = 0.1, member=None):
        """Write the seeded start on the device: member 0 with the counters
        reset (lg_seed_state), or one member of a batch (lg_seed_member)."""
        if member is None:
            check(lib().lg_seed_state(self._h, seed, noise), "seed_state")
        else:
            check(lib().lg_seed_member(self._h, member, seed, noise), "seed_member")

    def set_member(self, member: int, zo: float, mean_p_force_x: float,
                   mean_p_force_y: float = 0.0):
        check(lib().lg_member_params(self._h, member, zo, mean_p_force_x, mean_p_force_y),
              "member_params")

    def set_active(self, n: int):
        check(lib().lg_set_active(self._h, n), "set_active")

    def batch_info(self):
        a, b = C.c_int(), C.c_int()
        check(lib().lg_batch_info(self._h, C.byref(a), C.byref(b)), "batch_info")
        return a.value, b.value

    def copy_member(self, src: int, dst: int):
        check(lib().lg_copy_member(self._h, src, dst), "copy_member")
        self.nturb[dst] = self.nturb[src]

    def set_counters(self, jt: int, jt_total: int):
        check(lib().lg_set_counters(self._h, jt, jt_total), "set_counters")

    def counters(self):
        a, b = C.c_int(), C.c_int()
        check(lib().lg_get_counters(self._h, C.byref(a), C.byref(b)), "get_counters")
        return a.value, b.value

    def synchronize(self):
        check(lib().lg_synchronize(self._h), "synchronize")

    def resync(self):
        """Make the physical fields the state (the next step starts from them,
        as a restart from files does)."""
        check(lib().lg_resync(self._h), "resync")

    def upload_state(self, fields: dict, jt: int = 0, jt_total: int = 0):
        for n, a in fields.items():
            if n in FIELD_ID:
                self.upload(n, a)
        self.set_counters(jt, jt_total)

    # -- replaced subroutines ------------------------------------------------
    def filt_da(self, f, dfdx, dfdy):
        check(lib().lg_filt_da(self._h, _fid(f), _fid(dfdx), _fid(dfdy)), "filt_da")

    def ddx(self, f, dfdx):
        check(lib().lg_ddx(self._h, _fid(f), _fid(dfdx)), "ddx")

    def ddy(self, f, dfdy):
        check(lib().lg_ddy(self._h, _fid(f), _fid(dfdy)), "ddy")

    def ddxy(self, f, dfdx, dfdy):
        check(lib().lg_ddxy(self._h, _fid(f), _fid(dfdx), _fid(dfdy)), "ddxy")

    def test_filter(self, plane, width: int = 2):
        """test_filter (width 2) or test_test_filter (width 4) of one
        (ny, ld) plane; returns the filtered plane."""
        a = np.array(plane, dtype=np.float64, copy=True, order="C")
        if a.shape != (self.ny, self.ld):
            raise ValueError(f"expected a ({self.ny}, {self.ld}) plane, got {a.shape}")
        check(lib().lg_test_filter(self._h, width, a.ctypes.data_as(_lib._D)), "test_filter")
        return a

    def ddz_uv(self, f, dfdz):
        check(lib().lg_ddz_uv(self._h, _fid(f), _fid(dfdz)), "ddz_uv")

    def ddz_w(self, f, dfdz):
        check(lib().lg_ddz_w(self._h, _fid(f), _fid(dfdz)), "ddz_w")

    def wallstress(self):
        check(lib().lg_wallstress(self._h), "wallstress")

    def sgs_stag(self):
        check(lib().lg_sgs_stag(self._h), "sgs_stag")

    def lagrange_Sdep(self):
        check(lib().lg_lagrange_sdep(self._h), "lagrange_Sdep")

    def interpolag_Sdep(self):
        check(lib().lg_interpolag_sdep(self._h), "interpolag_Sdep")

    def divstress_uv(self):
        check(lib().lg_divstress_uv(self._h), "divstress_uv")

    def divstress_w(self):
        check(lib().lg_divstress_w(self._h), "divstress_w")

    def convec(self):
        check(lib().lg_convec(self._h), "convec")

    def press_stag_array(self):
        check(lib().lg_press_stag_array(self._h), "press_stag_array")

    def project(self):
        check(lib().lg_project(self._h), "project")

    def copy_field(self, src, dst):
        check(lib().lg_copy_field(self._h, _fid(src), _fid(dst)), "copy_field")

    def rhs_ab2(self):
        """RHS assembly, forcing and AB2 (the inline block of main.f90:211-308)."""
        check(lib().lg_rhs_ab2(self._h), "rhs_ab2")

    def rhs_gradp(self):
        check(lib().lg_rhs_gradp(self._h), "rhs_gradp")

    def step(self, nsteps: int = 1):
        check(lib().lg_step(self._h, nsteps), "step")

    def step_graph(self, nsteps: int = 1):
        """The same steps replayed from captured HIP graphs (lg_step_graph)."""
        check(lib().lg_step_graph(self._h, nsteps), "step_graph")

    def autotune(self, reps: int = 5) -> int:
        n = C.c_int()
        check(lib().lg_autotune(self._h, reps, C.byref(n)), "autotune")
        return n.value

    # -- diagnostics --------------------------------------------------------
    def _scalar(self, fn, member, what):
        out = C.c_double()
        check(fn(self._h, member, C.byref(out)), what)
        return out.value

    def rmsdiv(self, member: int = 0):
        return self._scalar(lib().lg_rmsdiv_member, member, "rmsdiv")

    def energy(self, member: int = 0):
        return self._scalar(lib().lg_energy_member, member, "energy")

    def max_cfl(self, member: int = 0):
        return self._scalar(lib().lg_max_cfl_member, member, "max_cfl")

    def tau_wall_bot(self, member: int = 0):
        return self._scalar(lib().lg_tau_wall_bot_member, member, "tau_wall_bot")

    def clock(self, member: int = 0) -> dict:
        """dt, tadv1, tadv2, lagran_dt and total_time of a member after its
        last step (with use_cfl_dt each member has its own)."""
        out = (C.c_double * 5)()
        check(lib().lg_clock_member(self._h, member, out), "clock")
        return dict(zip(("dt", "tadv1", "tadv2", "lagran_dt", "total_time"), list(out)))

    def set_clock(self, dt: float, total_time: float, fresh: bool = False, member: int = 0):
        """Restart a member's clock (total_time.dat): with use_cfl_dt, dt is
        the previous step's and fresh makes the next step an Euler step."""
        check(lib().lg_set_clock_member(self._h, member, dt, total_time, int(fresh)),
              "set_clock")

    # -- running time averages (time_average.f90) and spectra ---------------
    def stats_enable(self, nstart: int, nend: int, nskip: int = 1):
        check(lib().lg_stats_enable(self._h, nstart, nend, nskip), "stats_enable")

    def stats_info(self):
        t, i = C.c_double(), C.c_int()
        check(lib().lg_stats_info(self._h, C.byref(t), C.byref(i)), "stats_info")
        return t.value, bool(i.value)

    def stats_time(self, member: int = 0) -> float:
        """The averaging time of a member (each member's own with use_cfl_dt)."""
        t = C.c_double()
        check(lib().lg_stats_time(self._h, member, C.byref(t)), "stats_time")
        return t.value

    def stats_mean(self, name: str, member: int = 0):
        """Time average of a tavg quantity, [nz+1, ny, nx] float64."""
        a = np.empty((self.nz + 1, self.ny, self.cfg.nx))
        check(lib().lg_stats_mean(self._h, member, _lib.TAVG_NAMES.index(name),
                                  a.ctypes.data_as(_lib._D)), "stats_mean")
        return a

    def stats_sums(self, name: str, member: int = 0):
        """The raw time-weighted sum behind ``stats_mean`` (tavg.out's array)."""
        a = np.empty((self.nz + 1, self.ny, self.cfg.nx))
        check(lib().lg_stats_sums(self._h, member, _lib.TAVG_NAMES.index(name),
                                  a.ctypes.data_as(_lib._D)), "stats_sums")
        return a

    def stats_restore(self, total_time: float, sums: dict, member: int = 0):
        """Continue a previous run's sums (tavg%init reading tavg.out): call
        after ``stats_enable``; quantities missing from ``sums`` start at zero."""
        shp = (self.nz + 1, self.ny, self.cfg.nx)
        buf = np.zeros((len(_lib.TAVG_NAMES),) + shp)
        for q, n in enumerate(_lib.TAVG_NAMES):
            if n in sums:
                buf[q] = np.asarray(sums[n], dtype=np.float64).reshape(shp)
        check(lib().lg_stats_restore(self._h, member, float(total_time),
                                     buf.ctypes.data_as(_lib._D)), "stats_restore")

    def stats_profile(self, name: str, member: int = 0):
        return self.stats_mean(name, member).mean(axis=(1, 2))

    def stats_spectrum(self, component: int, member: int = 0):
        """Streamwise spectrum E(kx, z) of u (0), v (1) or w (2), [nz+1, nx/2]."""
        a = np.empty((self.nz + 1, self.cfg.nx // 2))
        check(lib().lg_stats_spectrum(self._h, member, component, a.ctypes.data_as(_lib._D)),
              "stats_spectrum")
        return a

    def diff_l2(self, other: "Case", field: str, k0: int, k1: int, member: int = 0,
                other_member: int = 0):
        """Relative L2 of this case's field against other's (x < nx, k0..k1)."""
        n, d = C.c_double(), C.c_double()
        check(lib().lg_diff_l2(self._h, member, other._h, other_member, _fid(field), k0, k1,
                               C.byref(n), C.byref(d)), "diff_l2")
        return math.sqrt(n.value / d.value) if d.value > 0 else math.sqrt(n.value)

    # -- per-kernel timing --------------------------------------------------
    def timing(self, on: bool):
        check(lib().lg_timing_enable(self._h, int(on)), "timing_enable")

    def timing_report(self):
        """{kind: (ms_total, launches, algorithmic_bytes)} for the timed launches."""
        out = {}
        for i, name in enumerate(_lib.KERNEL_KINDS):
            a, b, e = C.c_double(), C.c_double(), C.c_double()
            check(lib().lg_timing_report(self._h, i, C.byref(a), C.byref(b), C.byref(e)),
                  "timing_report")
            out[name] = (a.value, b.value, e.value)
        return out

    # -- actuator disks -----------------------------------------------------
    def turbines_init(self, specs, alpha1=1.5, alpha2=1.5, filter_cutoff=1e-2,
                      adm_correction=True, member: int = 0):
        """specs: iterable of (xloc, yloc, height, dia, thk, theta1, theta2, Ct');
        replaces the member's disks (an empty list removes them)."""
        a = np.ascontiguousarray(np.asarray(specs, dtype=np.float64).reshape(-1, 8))
        check(lib().lg_turbines_init_member(self._h, member, a.shape[0],
                                            a.ctypes.data_as(_lib._D), alpha1, alpha2,
                                            filter_cutoff, int(adm_correction)),
              "turbines_init")
        self.nturb[member] = a.shape[0]

    def turbines_forcing(self):
        check(lib().lg_turbines_forcing(self._h), "turbines_forcing")

    def turbines_state(self, member: int = 0):
        n = self.nturb[member]
        ud = np.empty(n)
        fn = np.empty(n)
        check(lib().lg_turbines_get_member(self._h, member, ud.ctypes.data_as(_lib._D),
                                           fn.ctypes.data_as(_lib._D), n), "turbines_get")
        return ud, fn

    def turbines_history(self, member: int = 0):
        """[n, 5]: u, v, w_uv at each disk's centre cell, u_d and u_d_T, as
        the last forcing saw them (turbines.f90:538-547, 589-593)."""
        n = self.nturb[member]
        out = np.empty((n, 5))
        check(lib().lg_turbines_history_member(self._h, member, out.ctypes.data_as(_lib._D), n),
              "turbines_history")
        return out

    def turbines_time_filter(self, T_avg_dim: float, z_i: float = 1.0, u_star: float = 1.0):
        """u_d_T filter of every disk (turbines.f90:561-566); T_avg_dim <= 0 is
        none."""
        check(lib().lg_turbines_time_filter(self._h, T_avg_dim, z_i, u_star), "time_filter")

    def turbines_set_ct(self, ct_prime, member: int = 0):
        """Dynamic Ct' of a member's disks (turbines.f90:504-515)."""
        a = np.ascontiguousarray(np.asarray(ct_prime, dtype=np.float64).ravel())
        check(lib().lg_turbines_set_ct_member(self._h, member, a.ctypes.data_as(_lib._D),
                                              a.size), "turbines_set_ct")

    def turbines_set_udt(self, u_d_T, member: int = 0):
        """u_d_T of a member's disks (u_d_T.dat, turbines.f90:229-250)."""
        a = np.ascontiguousarray(np.asarray(u_d_T, dtype=np.float64).ravel())
        check(lib().lg_turbines_set_udt_member(self._h, member, a.ctypes.data_as(_lib._D),
                                               a.size), "turbines_set_udt")


def tune_get(kind: str, precision: str, N: int, other: int):
    """(n1, n2, per_wg, variant) launch choice the library uses for a pass."""
    a, b, s, v = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    check(lib().lg_tune_get(_lib.PASS_KINDS.index(kind), LG_FP64 if precision == "fp64"
                            else LG_MIXED, N, other, C.byref(a), C.byref(b), C.byref(s),
                            C.byref(v)), "tune_get")
    return a.value, b.value, s.value, v.value


def tune_set(kind: str, precision: str, N: int, other: int, n1: int, n2: int, per_wg: int,
             variant: int = 0):
    check(lib().lg_tune_set(_lib.PASS_KINDS.index(kind), LG_FP64 if precision == "fp64"
                            else LG_MIXED, N, other, n1, n2, per_wg, variant), "tune_set")


def tune_clear():
    check(lib().lg_tune_clear(), "tune_clear")


def tune_table():
    """The tuned launch choices as a list of dicts (JSON-ready)."""
    n = C.c_int()
    check(lib().lg_tune_count(C.byref(n)), "tune_count")
    out = []
    for i in range(n.value):
        v = [C.c_int() for _ in range(8)]
        check(lib().lg_tune_entry(i, *[C.byref(x) for x in v]), "tune_entry")
        k, pr, N, o, a, b, s, var = (x.value for x in v)
        out.append({"kind": _lib.PASS_KINDS[k], "precision": "fp64" if pr == LG_FP64 else "mixed",
                    "N": N, "other": o, "n1": a, "n2": b, "per_wg": s, "variant": var})
    return out


def tune_save(path: str):
    import json
    with open(path, "w") as f:
        json.dump(tune_table(), f, indent=1)


def tune_load(path: str) -> int:
    """Load choices saved by tune_save; returns how many were set."""
    import json
    with open(path) as f:
        rows = json.load(f)
    for r in rows:
        tune_set(r["kind"], r["precision"], r["N"], r["other"], r["n1"], r["n2"], r["per_wg"],
                 r.get("variant", 0))
    return len(rows)


def seeded_state(cfg: Config, seed: int, noise: float = 0.1):
    """The seeded vel.out-equivalent start (the ic_file path,
    initial.f90:226-239) as host arrays in the reference layout: log-law u
    plus counter-based uniform noise (splitmix64 keyed by seed, field,
    i, j, k), w = 0 at the wall and lid, Cs_opt2 = 0.03, F_* = 0."""
    nx, ny, nz = cfg.nx, cfg.ny, cfg.Nz + 1
    ld = 2 * (nx // 2 + 1)
    dz = cfg.lz / (nz - 1)
    shp = (nz + 1, ny, ld)
    out = {n: np.zeros(shp) for n in ("u", "v", "w", "Cs_opt2")}
    k = np.arange(1, nz)[:, None, None]
    j = np.arange(ny)[None, :, None]
    i = np.arange(nx)[None, None, :]
    zk = (k - 0.5) * dz
    ubar = (1.0 / cfg.vonk) * np.log(zk / cfg.zo)
    amp = noise * ubar
    for fid, name in enumerate(("u", "v", "w")):
        r = uniform01(seed, fid, i, j, k)
        base = ubar if name == "u" else 0.0
        out[name][1:nz, :, :nx] = base + amp * (2.0 * r - 1.0)
    out["w"][1] = 0.0
    out["w"][nz] = 0.0
    out["u"][nz] = out["u"][nz - 1]
    out["v"][nz] = out["v"][nz - 1]
    out["Cs_opt2"][1:nz + 1] = 0.03
    for name in ("u", "v", "w"):
        out[name][0] = -1234567890.0
    return out


_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def uniform01(seed, fid, i, j, k):
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0x100000001B3)) & _M64
        key = _splitmix64(key ^ np.uint64(fid))
        key = _splitmix64(key ^ (np.asarray(k, np.uint64) << np.uint64(42)
                                 ^ np.asarray(j, np.uint64) << np.uint64(21)
                                 ^ np.asarray(i, np.uint64)))
    return (key >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def aligned_farm(cfg: Config, nxt=4, nyt=4, dia=0.1, hub=0.1, sx_d=7.0, sy_d=5.0,
                 ct_prime=1.33, thk=0.0):
    """Regular aligned actuator-disk array centred in the domain
    (x spacing sx_d D, y spacing sy_d D), rows of (x, y, h, D, thk, th1, th2, Ct')."""
    x0 = 0.5 * cfg.lx - 0.5 * (nxt - 1) * sx_d * dia
    y0 = 0.5 * cfg.ly - 0.5 * (nyt - 1) * sy_d * dia
    return [(x0 + i * sx_d * dia, y0 + j * sy_d * dia, hub, dia, thk, 0.0, 0.0, ct_prime)
            for i in range(nxt) for j in range(nyt)]
