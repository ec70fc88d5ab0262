"""A lesgo run from ``lesgo.conf`` on the GPU: ``python -m lesgo_amd.run [path]``.

The control flow of main.f90 (initialize, time loop 130-468, finalize) for
the configurations this time step implements (``LesgoConf.check_supported``),
reading and writing the reference's files in ``path``:

- input: ``lesgo.conf``; ``vel.out`` + ``total_time.dat`` when present
  (initu = .true., initial.f90:96-110, io.f90:62-95); ``input_turbines/
  param.dat`` (read_param), ``input_turbines/{theta1,theta2,Ct_prime}.dat``
  (dynamic yaw / tilt / thrust schedules) and ``turbine/u_d_T.dat``;
- output: ``output/lesgo_param.out`` (param_output.f90); the iteration block
  every ``wbase`` steps (main.f90:395-436); ``output/check_ke.dat`` every
  ``nenergy`` steps (io.f90:99-141); ``vel.out``, ``grid.out``,
  ``total_time.dat`` every ``checkpoint_nskip`` steps with
  ``checkpoint_data`` and at the end (io.f90:1173-1267); ``output/vel.<jt>.c0.bin``,
  ``output/vort.<jt>.c0.bin`` and ``output/pres.<jt>.c0.bin`` with
  ``domain_calc`` (io.f90:786-884, one process = coord 0); point and x/y/z
  plane samples with ``point_calc`` / ``*plane_calc`` (io.f90:585-676, the
  writers of ``lesgo_amd.samples``); the time averages ``output/{veluv,velw,vel2,tau,pres,
  force,cs_opt2,vort}_avg.bin`` and ``output/rs.bin`` with ``tavg_calc``
  (time_average.f90:322-655) and their restart file ``tavg.out``
  (time_average.f90:665-716, read back on a restart as tavg%init does);
  ``turbine/u_d_T.dat`` (turbines.f90:657-679)
  and the per-turbine histories ``turbine/turbine_<s>.dat`` every ``tbase``
  steps (turbines.f90:589-593).

A new run starts from the seeded log-law field (``--seed``) in place of the
reference's /dev/urandom ic_les.
Steps replay captured graphs unless a dynamic turbine schedule changes the
disks between steps; dynamic schedules need a static dt (their time is the
dimensional time at the end of the step, which with use_cfl_dt is only known
on the device).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from .case import Case
from .checkpoint import (read_checkpoint, read_tavg_checkpoint, write_checkpoint,
                         write_tavg_checkpoint)
from .conf import (ConfError, LesgoConf, linear_interp, place_turbines, read_conf,
                   read_control_files, write_param_out)
from .samples import Sampler, vorticity


def _write_direct(fname, arrays, nz):
    """Direct-access records of nx*ny*nz doubles (levels 1..nz, x fastest)."""
    with open(fname, "wb") as f:
        for a in arrays:
            f.write(np.ascontiguousarray(a[1:nz + 1]).astype("<f8").tobytes())


def write_tavg(case: Case, c: LesgoConf, member: int = 0):
    """time_average.f90:322-655 binary branch: the means, then the Reynolds
    stresses from them."""
    out = os.path.join(c.path, "output")
    nz = case.nz
    m = {n: case.stats_mean(n, member) for n in
         ("u", "v", "w", "w_uv", "u_w", "v_w", "u2", "v2", "w2", "uv", "uw", "vw", "txx",
          "tyy", "tzz", "txy", "txz", "tyz", "p", "fx", "fy", "fz", "cs_opt2", "vortx",
          "vorty", "vortz")}
    _write_direct(os.path.join(out, "veluv_avg.bin"), [m["u"], m["v"], m["w_uv"]], nz)
    _write_direct(os.path.join(out, "velw_avg.bin"), [m["w"]], nz)
    _write_direct(os.path.join(out, "vel2_avg.bin"),
                  [m["u2"], m["v2"], m["w2"], m["uw"], m["vw"], m["uv"]], nz)
    _write_direct(os.path.join(out, "tau_avg.bin"),
                  [m["txx"], m["txy"], m["tyy"], m["txz"], m["tyz"], m["tzz"]], nz)
    _write_direct(os.path.join(out, "pres_avg.bin"), [m["p"]], nz)
    if c.turbines:
        _write_direct(os.path.join(out, "force_avg.bin"), [m["fx"], m["fy"], m["fz"]], nz)
    _write_direct(os.path.join(out, "cs_opt2.bin"), [m["cs_opt2"]], nz)
    _write_direct(os.path.join(out, "vort_avg.bin"), [m["vortx"], m["vorty"], m["vortz"]], nz)
    rs = [m["u2"] - m["u"] * m["u"], m["v2"] - m["v"] * m["v"], m["w2"] - m["w"] * m["w"],
          m["uw"] - m["u_w"] * m["w"], m["vw"] - m["v_w"] * m["w"], m["uv"] - m["u"] * m["v"]]
    _write_direct(os.path.join(out, "rs.bin"), rs, nz)


def write_domain(case: Case, c: LesgoConf, jt: int, member: int = 0):
    """io.f90:786-884 binary branch: u, v, w interpolated to the uv grid;
    the vorticity on the w grid from the derivative fields (io.f90:823-858);
    p - (u^2 + v^2 + w_uv^2)/2 (levels 1..nz, i < nx)."""
    nz, nx = case.nz, c.nx
    u = case.download("u", member)[:, :, :nx]
    v = case.download("v", member)[:, :, :nx]
    w = case.download("w", member)[:, :, :nx]
    p = case.download("p", member)[:, :, :nx]
    w_uv = np.empty_like(w)
    w_uv[1:nz] = 0.5 * (w[2:nz + 1] + w[1:nz])
    w_uv[nz] = w_uv[nz - 1]
    pr = p - 0.5 * (u * u + w_uv * w_uv + v * v)
    der = {n: case.download(n, member)[:, :, :nx]
           for n in ("dudy", "dudz", "dvdx", "dvdz", "dwdx", "dwdy")}
    out = os.path.join(c.path, "output")
    _write_direct(os.path.join(out, f"vel.{jt}.c0.bin"), [u, v, w_uv], nz)
    _write_direct(os.path.join(out, f"vort.{jt}.c0.bin"), list(vorticity(der, nz)), nz)
    _write_direct(os.path.join(out, f"pres.{jt}.c0.bin"), [pr], nz)


def write_samples(case: Case, sampler: Sampler, jt: int, path: str, member: int = 0):
    """The point and plane samples due at step jt (output_loop, io.f90:585-676)."""
    due = sampler.due(jt)
    if not any(due.values()):
        return
    u, v, w = (case.download(n, member) for n in ("u", "v", "w"))
    sampler.write(u, v, w, jt, case.clock(member)["total_time"], os.path.join(path, "output"),
                  due)


class Runner:
    """One lesgo run on the GPU; ``run()`` is the time loop."""

    def __init__(self, path: str = "./", precision: str = "fp64", seed: int = 0,
                 quiet: bool = False):
        self.path = path if path.endswith("/") else path + "/"
        self.quiet = quiet
        self.conf = c = read_conf(self.path)
        for msg in c.messages:
            self.say(msg)
        self.restart = os.path.exists(os.path.join(self.path, "vel.out"))
        c.initu = self.restart
        c.inilag = not self.restart
        self.cfg = c.to_config(precision, initu=self.restart)
        self.schedules = read_control_files(c) if c.turbines else {}
        if self.schedules and c.use_cfl_dt:
            raise ConfError("dynamic turbine schedules need use_cfl_dt = .false. here")
        self.case = Case(self.cfg)
        os.makedirs(os.path.join(self.path, "output"), exist_ok=True)
        self.sampler = Sampler(c)
        write_param_out(c)
        self.total_time_dim = 0.0
        if self.restart:
            self.say("--> Reading initial velocity field from file")
            jt, t = read_checkpoint(self.case, self.path)
            if not c.cumulative_time:
                self.case.set_counters(0, 0)
                self.case.set_clock(self.case.clock()["dt"], 0.0, c.use_cfl_dt)
            else:
                w = open(os.path.join(self.path, "total_time.dat")).read().split()
                self.total_time_dim = float(w[2])
        else:
            self.say("--> Creating initial velocity field from the seeded log law")
            self.case.seed_state(seed)
        self.specs = None
        if c.turbines:
            self.specs = [list(t) for t in place_turbines(c)]
            self._apply_schedules(self.total_time_dim)
            self.case.turbines_init(self.specs, c.alpha1, c.alpha2, c.filter_cutoff,
                                    c.adm_correction)
            self.case.turbines_time_filter(c.T_avg_dim, c.z_i, c.u_star)
            fu = os.path.join(self.path, "turbine", "u_d_T.dat")
            if os.path.exists(fu):
                vals = [float(x) for x in open(fu).read().split()]
                self.case.turbines_set_udt(vals[:len(self.specs)])
            # the per-turbine history files, opened for appending at
            # initialisation (turbines.f90:262-267)
            os.makedirs(os.path.join(self.path, "turbine"), exist_ok=True)
            for q in range(len(self.specs)):
                open(self._turbine_file(q), "a").close()
        if c.tavg_calc:
            self.case.stats_enable(c.tavg_nstart, c.tavg_nend, c.tavg_nskip)
            # earlier sums continue (tavg%init reads tavg.out when present)
            if read_tavg_checkpoint(self.case, self.path):
                self.say("--> Continuing the time averages from tavg.out")

    def say(self, s):
        if not self.quiet:
            print(s, flush=True)

    def _turbine_file(self, q):
        return os.path.join(self.path, "turbine", f"turbine_{q + 1}.dat")

    def write_turbine_history(self):
        """One record per disk (turbines.f90:589-593): total_time_dim, u, v,
        w_uv at the disk centre, -u_d, -u_d_T, theta1, theta2, Ct_prime of the
        step's forcing.  With use_cfl_dt the dimensional time is the device
        clock's total time times z_i / u_star."""
        c = self.conf
        t_dim = (self.case.clock()["total_time"] * c.z_i / c.u_star if c.use_cfl_dt
                 else self.total_time_dim)
        h = self.case.turbines_history()
        for q, sp in enumerate(self.specs):
            vals = [t_dim, h[q, 0], h[q, 1], h[q, 2], -h[q, 3], -h[q, 4], sp[5], sp[6], sp[7]]
            with open(self._turbine_file(q), "a") as f:
                f.write(" " + " ".join(repr(float(x)) for x in vals) + "\n")

    def _apply_schedules(self, t_dim) -> str:
        """Interpolate the dynamic schedules at t_dim into self.specs; returns
        "" (no change), "ct" or "nodes" (yaw / tilt changed)."""
        what = ""
        for name, col in (("theta1", 5), ("theta2", 6), ("Ct_prime", 7)):
            if name not in self.schedules:
                continue
            times, vals = self.schedules[name]
            for s in range(len(self.specs)):
                v = linear_interp(times, vals[s], t_dim)
                if v != self.specs[s][col]:
                    self.specs[s][col] = v
                    what = "nodes" if col < 7 else (what or "ct")
        return what

    def _step(self):
        c = self.conf
        if not c.use_cfl_dt:
            # total_time_dim += dt_dim (main.f90:150); the forcing of this
            # step sees the time at its end
            self.total_time_dim = self.total_time_dim + c.dt_dim
        if self.schedules:
            what = self._apply_schedules(self.total_time_dim)
            if what == "nodes":
                udt, _ = self.case.turbines_state()
                self.case.turbines_init(self.specs, c.alpha1, c.alpha2, c.filter_cutoff,
                                        c.adm_correction)
                self.case.turbines_set_udt(udt)
            elif what == "ct":
                self.case.turbines_set_ct([s[7] for s in self.specs])
            self.case.step(1)
        else:
            self.case.step_graph(1)

    def _report(self, jt, wall):
        c = self.case
        ck = c.clock()
        self.say("=" * 51)
        self.say(f"  Wall time per step (s): {wall:15.7E}")
        self.say(f"  Iteration: {jt:9d}")
        self.say(f"  Time step: {ck['dt']:15.7E}")
        self.say(f"  CFL: {c.max_cfl():15.7E}")
        self.say(f"  AB2 TADV1, TADV2: {ck['tadv1']:15.7E}{ck['tadv2']:15.7E}")
        self.say(f"  Velocity divergence metric: {c.rmsdiv():15.7E}")
        self.say(f"  Kinetic energy: {c.energy():15.7E}")
        self.say(f"  Bot wall stress: {c.tau_wall_bot():15.7E}")

    def run(self):
        c, case = self.conf, self.case
        jt0 = case.counters()[1]
        t0 = time.perf_counter()
        tlast, jlast = t0, jt0
        for jt in range(jt0 + 1, c.nsteps + 1):
            self._step()
            if c.runtime > 0 and time.perf_counter() - t0 > c.runtime:
                self.say("Specified runtime exceeded. Exiting simulation.")
                break
            if jt % c.nenergy == 0:
                ke = case.energy()
                with open(os.path.join(self.path, "output", "check_ke.dat"), "a") as f:
                    f.write(f" {case.clock()['total_time']!r} {ke!r}\n")
            if jt % c.wbase == 0:
                case.synchronize()
                now = time.perf_counter()
                self._report(jt, (now - tlast) / max(jt - jlast, 1))
                tlast, jlast = now, jt
            if (c.domain_calc and c.domain_nstart <= jt <= c.domain_nend
                    and (jt - c.domain_nstart) % c.domain_nskip == 0):
                write_domain(case, c, jt)
            write_samples(case, self.sampler, jt, self.path)
            if self.specs and jt % c.tbase == 0:
                self.write_turbine_history()
            if c.checkpoint_data and jt % c.checkpoint_nskip == 0:
                self.finalize_files()
        case.synchronize()
        wall = time.perf_counter() - t0
        self.finalize_files()
        if c.tavg_calc and case.stats_info()[1]:
            write_tavg(case, c)
        self.say(f"Simulation wall time (s) : {wall:15.7E}")
        self.say("Simulation complete")
        return wall

    def finalize_files(self):
        c = self.conf
        if c.use_cfl_dt:
            # dt_dim = dt z_i / u_star summed: the device clock's time scaled
            self.total_time_dim = self.case.clock()["total_time"] * c.z_i / c.u_star
        write_checkpoint(self.case, self.path, total_time_dim=self.total_time_dim)
        if c.tavg_calc:
            write_tavg_checkpoint(self.case, self.path)      # io.f90:1229
        if self.specs:
            os.makedirs(os.path.join(self.path, "turbine"), exist_ok=True)
            udt, _ = self.case.turbines_state()
            with open(os.path.join(self.path, "turbine", "u_d_T.dat"), "w") as f:
                for v in udt:
                    f.write(f" {float(v)!r}\n")

    def close(self):
        self.case.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("path", nargs="?", default="./")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    try:
        r = Runner(a.path, a.precision, a.seed)
    except ConfError as e:
        print(f"lesgo.conf: {e}", file=sys.stderr)
        return 1
    try:
        r.run()
    finally:
        r.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
