"""Serving loop driver: a Poisson stream of LES jobs through the library's
native scheduler (lg_server_*, continuous batching under an HBM budget).

    python -m lesgo_amd.serve --minutes 10 --rate 2 --budget-gb 200

Each job's grid is drawn uniformly from {64^3, 96x48x64, 128x64x96, 128^3}
(Lx = 2 pi, Ly = pi, Lz = 1), its step count uniformly from [200, 2000]; it
has its own seed, z0 log-uniform in [1e-5, 1e-3], and a 30 % chance of a
2x2 aligned actuator-disk array (D = 0.15, hub 0.1, 7D x 5D, Ct' = 1.33).
Jobs run in mixed precision.  Prints one JSON line with p50/p99 job
completion latency, p50/p99 per-step (tick) latency and aggregate
throughput; with --serial beside the same arrival process served one job at
a time in plain FP64 (measured), with --saturation N the sustained job rate
of the first N jobs submitted at once.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import _lib
from ._lib import LG_FP64, check, lib
from .case import Case, Config, aligned_farm

GRIDS = ((64, 64, 64), (96, 48, 64), (128, 64, 96), (128, 128, 128))
DT = 2e-4         # fresh starts on all four grids stay bounded at this step


class Server:
    """Native scheduler handle (lg_server_*)."""

    def __init__(self, budget_bytes: float, pool_members: int = 16):
        h = C.c_void_p()
        check(lib().lg_server_create(budget_bytes, pool_members, C.byref(h)), "server_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().lg_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def now(self) -> float:
        t = C.c_double()
        check(lib().lg_server_now(self._h, C.byref(t)), "server_now")
        return t.value

    def submit(self, cfg: Config, seed: int, nsteps: int, disks=(), t_arrive=None,
               keep=False, noise: float = 0.1) -> int:
        a = np.ascontiguousarray(np.asarray(disks, dtype=np.float64).reshape(-1, 8))
        j = C.c_int()
        check(lib().lg_server_submit(self._h, C.byref(cfg.to_c()), seed, noise, nsteps,
                                     a.shape[0], a.ctypes.data_as(_lib._D),
                                     self.now() if t_arrive is None else t_arrive, int(keep),
                                     C.byref(j)), "server_submit")
        return j.value

    def tick(self):
        t, r, q = C.c_double(), C.c_int(), C.c_int()
        check(lib().lg_server_tick(self._h, C.byref(t), C.byref(r), C.byref(q)), "server_tick")
        return t.value, r.value, q.value

    def job(self, j: int) -> dict:
        st, sd = C.c_int(), C.c_int()
        ta, tb, tc = C.c_double(), C.c_double(), C.c_double()
        check(lib().lg_server_job(self._h, j, C.byref(st), C.byref(sd), C.byref(ta),
                                  C.byref(tb), C.byref(tc)), "server_job")
        return {"state": st.value, "steps_done": sd.value, "t_arrive": ta.value,
                "t_admit": tb.value, "t_done": tc.value}

    def result(self, j: int, field: str, shape):
        a = np.empty(shape)
        check(lib().lg_server_result(self._h, j, _lib.FIELD_ID[field],
                                     a.ctypes.data_as(_lib._D)), "server_result")
        return a

    def stats(self) -> dict:
        ps, nb = C.c_double(), C.c_double()
        tk, npl = C.c_long(), C.c_int()
        check(lib().lg_server_stats(self._h, C.byref(ps), C.byref(tk), C.byref(nb),
                                    C.byref(npl)), "server_stats")
        return {"pts_steps": ps.value, "ticks": tk.value, "bytes_used": nb.value,
                "pools": npl.value}


def job_stream(minutes: float, rate: float, seed: int = 0, steps=(200, 2000), grids=GRIDS,
               p_disks: float = 0.3):
    """Arrival times (Poisson, `rate` jobs/s) and per-job parameters."""
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > 60.0 * minutes:
            break
        g = grids[rng.integers(len(grids))]
        jobs.append({
            "t": t, "grid": tuple(int(x) for x in g),
            "nsteps": int(rng.integers(steps[0], steps[1] + 1)),
            "seed": int(rng.integers(1 << 30)),
            "zo": float(10.0 ** rng.uniform(-5.0, -3.0)),
            "disks": bool(rng.random() < p_disks),
        })
    return jobs


def job_config(job, precision="mixed", dt=DT) -> Config:
    """A job is a fresh start: Euler first step, Smagorinsky until DYN_init,
    then the Lagrangian dynamic model."""
    nx, ny, Nz = job["grid"]
    return Config(nx, ny, Nz, dt=dt, zo=job["zo"], precision=precision, initu=False)


def job_disks(cfg: Config, on: bool):
    if not on:
        return []
    return aligned_farm(cfg, 2, 2, dia=0.15, hub=0.1, sx_d=7.0, sy_d=5.0, ct_prime=1.33)


def run_stream(jobs, budget_gb=200.0, pool_members=16, progress=None, drain=True):
    """Feed `jobs` at their arrival times, tick until all are done (or, with
    drain=False, until the last arrival); returns the summary dict."""
    with Server(budget_gb * 1e9, pool_members) as srv:
        ids, ticks = [], []
        nxt, t_last = 0, -1e9
        t_start = srv.now()
        while True:
            now = srv.now()
            while nxt < len(jobs) and jobs[nxt]["t"] <= now - t_start:
                j = jobs[nxt]
                cfg = job_config(j)
                ids.append(srv.submit(cfg, j["seed"], j["nsteps"], job_disks(cfg, j["disks"]),
                                      t_arrive=t_start + j["t"]))
                nxt += 1
            secs, running, queued = srv.tick()
            if running or queued:
                ticks.append(secs)
            if progress and srv.now() - t_last > 20.0:
                t_last = srv.now()
                progress(f"[serve] t={t_last - t_start:.0f}s submitted={nxt}/{len(jobs)} "
                         f"running={running} queued={queued} ticks={len(ticks)}")
            if nxt == len(jobs) and running == 0 and queued == 0:
                break
            if not drain and nxt == len(jobs):
                break
            if running == 0 and queued == 0 and nxt < len(jobs):
                time.sleep(min(0.01, max(0.0, jobs[nxt]["t"] - (srv.now() - t_start))))
        wall = srv.now() - t_start
        st = srv.stats()
        info = [srv.job(i) for i in ids]
    done = [x for x in info if x["state"] == 2]
    lat = np.array([x["t_done"] - x["t_arrive"] for x in done])
    tk = np.array(ticks) * 1e3
    return {
        "jobs": len(jobs), "completed": len(done), "wall_s": wall,
        "job_latency_s_p50": float(np.percentile(lat, 50)) if len(lat) else None,
        "job_latency_s_p99": float(np.percentile(lat, 99)) if len(lat) else None,
        "step_ms_p50": float(np.percentile(tk, 50)) if len(tk) else None,
        "step_ms_p99": float(np.percentile(tk, 99)) if len(tk) else None,
        "throughput_Mpts_steps_s": st["pts_steps"] / wall / 1e6,
        "pools_at_end": st["pools"], "ticks": st["ticks"],
        "queue_wait_s_p50": float(np.median([x["t_admit"] - x["t_arrive"] for x in done]))
        if done else None,
    }


def run_serial_fp64(jobs, progress=None):
    """The same arrival process served one job at a time in plain FP64,
    first in first out: job i starts at max(its arrival, the previous job's
    completion) on a case of its own (created, seeded, graph-replayed for
    its steps, destroyed), as a host loop without batching would run it.
    Returns the latency distribution and the wall time, measured."""
    lat, busy = [], 0.0
    t_start = time.perf_counter()
    t_last = t_start
    for n, j in enumerate(jobs):
        now = time.perf_counter() - t_start
        if now < j["t"]:
            time.sleep(j["t"] - now)
        cfg = job_config(j, precision="fp64")
        a = time.perf_counter()
        with Case(cfg) as c:
            c.seed_state(j["seed"])
            disks = job_disks(cfg, j["disks"])
            if disks:
                c.turbines_init(disks)
            c.set_counters(0, 0)
            c.step_graph(j["nsteps"])
            c.synchronize()
        b = time.perf_counter()
        busy += b - a
        lat.append(b - t_start - j["t"])
        if progress and b - t_last > 20.0:
            t_last = b
            progress(f"[serial fp64] t={b - t_start:.0f}s done={n + 1}/{len(jobs)}")
    wall = time.perf_counter() - t_start
    pts = sum(j["grid"][0] * j["grid"][1] * j["grid"][2] * j["nsteps"] for j in jobs)
    lat = np.array(lat)
    return {"wall_s": wall, "busy_s": busy,
            "job_latency_s_p50": float(np.percentile(lat, 50)),
            "job_latency_s_p99": float(np.percentile(lat, 99)),
            "throughput_Mpts_steps_s": pts / wall / 1e6,
            "busy_Mpts_steps_s": pts / busy / 1e6}


def saturation(jobs, budget_gb=200.0, pool_members=16, progress=None):
    """Every job of `jobs` submitted at once: the server's sustained job
    rate (jobs/s) and throughput for this job mix, i.e. the arrival rate
    beyond which its queue grows without bound."""
    burst = [dict(j, t=0.0) for j in jobs]
    res = run_stream(burst, budget_gb, pool_members, progress=progress)
    return {"jobs": len(jobs), "wall_s": res["wall_s"],
            "jobs_per_s": len(jobs) / res["wall_s"],
            "throughput_Mpts_steps_s": res["throughput_Mpts_steps_s"],
            "step_ms_p50": res["step_ms_p50"], "step_ms_p99": res["step_ms_p99"]}


def fp64_serial_estimate(jobs):
    """Seconds the jobs would take one after another in plain FP64: per grid,
    the FP64 step time before DYN_init (Smagorinsky phase) and after it
    (dynamic model every cs_count steps), measured here on a seeded fresh
    start, times each job's steps in each phase."""
    per = {}
    for g in sorted({j["grid"] for j in jobs}):
        cfg = Config(*g, dt=DT, precision="fp64", initu=False)
        with Case(cfg) as c:
            c.seed_state(1)
            c.step_graph(5)
            c.synchronize()
            t0 = time.perf_counter()
            c.step_graph(10)
            c.synchronize()
            t_smag = (time.perf_counter() - t0) / 10
            c.step_graph(cfg.dyn_init - 15)
            c.synchronize()
            t0 = time.perf_counter()
            c.step_graph(20)
            c.synchronize()
            per[g] = (t_smag, (time.perf_counter() - t0) / 20, cfg.dyn_init)
    total = 0.0
    for j in jobs:
        ts, td, di = per[j["grid"]]
        total += ts * min(j["nsteps"], di - 1) + td * max(0, j["nsteps"] - di + 1)
    pts = sum(j["grid"][0] * j["grid"][1] * j["grid"][2] * j["nsteps"] for j in jobs)
    return {"seconds": total, "Mpts_steps_s": pts / total / 1e6,
            "fp64_step_ms": {"x".join(map(str, g)): [1e3 * v[0], 1e3 * v[1]]
                             for g, v in per.items()}}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--minutes", type=float, default=10.0)
    ap.add_argument("--rate", type=float, default=2.0)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--pool-members", type=int, default=16)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--max-steps", type=int, default=2000)
    ap.add_argument("--serial", action="store_true",
                    help="also replay the stream one job at a time in plain FP64 (measured)")
    ap.add_argument("--saturation", type=int, default=0, metavar="NJOBS",
                    help="also submit the first NJOBS jobs at once (sustained job rate)")
    args = ap.parse_args(argv)
    jobs = job_stream(args.minutes, args.rate, args.seed, steps=(200, args.max_steps))
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    log(f"[serve] {len(jobs)} jobs over {args.minutes} min at {args.rate}/s")
    res = run_stream(jobs, args.budget_gb, args.pool_members, progress=log)
    if args.serial:
        ser = run_serial_fp64(jobs, progress=log)
        res["fp64_serial"] = ser
        res["job_latency_p50_ratio_serial_over_served"] = (ser["job_latency_s_p50"]
                                                           / res["job_latency_s_p50"])
    else:
        ser = fp64_serial_estimate(jobs)
        res["fp64_serial_estimate_s"] = ser["seconds"]
        res["fp64_step_ms"] = ser["fp64_step_ms"]
    if args.saturation:
        res["saturation"] = saturation(jobs[:args.saturation], args.budget_gb,
                                       args.pool_members, progress=log)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
