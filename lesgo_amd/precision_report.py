"""Mixed-precision error report on the 128^3 neutral half channel.

    python -m lesgo_amd.precision_report [--steps 1000] [--seeds 0 1 2 3]

Runs the seeds as one FP64 batch and one mixed batch (fresh starts, dt =
2e-4), samples the per-step velocity error of every member against its FP64
twin (relative L2 over the valid levels, computed on the device), sums the
running time averages and streamwise spectra over the last half of the run,
and prints one JSON line: per seed the velocity error at a few steps and its
maximum, the divergence of both runs, and the relative L2 error of the mean
profiles <u>, <v>, the resolved variance profiles <u'u'>, <w'w'> and the
spectra E_uu, E_vv, E_ww over the averaging window.  The FP64 batch is the
reference: its step equals the oracle's to 1e-10 (the parity tests).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from .case import Case, Config


def rel(a, b):
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / nb if nb > 0 else float(np.linalg.norm(a))


def run(n=128, Nz=128, steps=1000, seeds=(0, 1, 2, 3), dt=2e-4, every=10, log=None):
    cfgs = {p: Config(n, n, Nz, dt=dt, precision=p, initu=False) for p in ("fp64", "mixed")}
    M = len(seeds)
    cases = {p: Case(cfgs[p], members=M) for p in cfgs}
    try:
        for c in cases.values():
            for m, sd in enumerate(seeds):
                c.seed_state(sd, 0.1, member=m)
            c.set_counters(0, 0)
            c.stats_enable(steps // 2, steps, 1)
        nz = cases["fp64"].nz
        curve = {m: [] for m in range(M)}
        t0 = time.perf_counter()
        for s in range(every, steps + 1, every):
            for c in cases.values():
                c.step_graph(every)
            for m in range(M):
                e = max(cases["mixed"].diff_l2(cases["fp64"], f, 1, nz - 1, m, m)
                        for f in ("u", "v", "w"))
                curve[m].append((s, e))
            if log and s % 200 == 0:
                log(f"[precision] step {s}: max vel err {max(curve[m][-1][1] for m in range(M)):.2e}")
        wall = time.perf_counter() - t0
        out = {"grid": [n, n, Nz], "steps": steps, "dt": dt, "seeds": list(seeds),
               "average_window": [steps // 2, steps], "wall_s": wall, "per_seed": []}
        for m, sd in enumerate(seeds):
            r = {"seed": sd}
            ce = dict(curve[m])
            r["vel_err_at"] = {str(k): ce[k] for k in (10, 100, steps // 2, steps) if k in ce}
            r["vel_err_max"] = max(e for _, e in curve[m])
            r["rmsdiv"] = {p: cases[p].rmsdiv(m) for p in cases}
            K = slice(1, nz)
            prof = {}
            for name in ("u", "v"):
                prof[name] = rel(cases["mixed"].stats_profile(name, m)[K],
                                 cases["fp64"].stats_profile(name, m)[K])
            for name, mean in (("u2", "u"), ("w2", "w")):
                var = {}
                for p in cases:
                    A = cases[p].stats_mean(name, m)
                    B = cases[p].stats_mean(mean, m)
                    var[p] = (A - B * B).mean(axis=(1, 2))[K]
                prof[name.replace("2", "'" + name[0] + "'")] = rel(var["mixed"], var["fp64"])
            r["profile_err"] = prof
            r["spectrum_err"] = {
                lab: rel(cases["mixed"].stats_spectrum(c, m)[K], cases["fp64"].stats_spectrum(c, m)[K])
                for c, lab in enumerate(("E_uu", "E_vv", "E_ww"))}
            out["per_seed"].append(r)
        return out
    finally:
        for c in cases.values():
            c.close()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--seeds", type=int, nargs="+", default=[0, 1, 2, 3])
    ap.add_argument("--n", type=int, default=128)
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    print(json.dumps(run(args.n, args.n, args.steps, tuple(args.seeds), log=log)), flush=True)


if __name__ == "__main__":
    main()
