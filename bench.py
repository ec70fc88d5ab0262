"""Throughput of the lesgo time step on one MI355X.

Workload (default): wind-farm LES, 256 x 256 x 192 grid on Lx = 2 pi,
Ly = pi, Lz = 1, 4 x 4 aligned actuator-disk array (D = 0.1, hub 0.1,
7D x 5D, Ct' = 1.33), Lagrangian scale-dependent SGS, equilibrium wall
model z0 = 1e-4, dt = 1e-4, seeded log-law start (seed 7) run as a fresh
start (Euler first step, Smagorinsky Cs^2 = 0.03 until DYN_init = 100, then
the dynamic model every 5 steps).  The FP64 and the mixed-precision case
run the first DYN_init steps untimed; their velocity difference is measured
on the device at that point; then each runs a warm-up and three passes of
the timed steps: graph-replayed with no per-launch events (`value`, the
headline), eager with no events, and eager with HIP events around every
launch of the transform and tridiagonal kernel classes (the per-kernel
roofline).  `value` is the mixed step's grid-pts * steps / s; the FP64
figure and the ratio are reported beside it, with the whole-step roofline
(algorithmic bytes and flops per step against the HBM and vector peaks).
The ensemble, serving and CPU legs follow.

Prints one JSON line.  `--gpus N` with N > 1 exits: the library runs a case
on one GPU, with no sharding.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpts·steps/s at 256³-class, mixed vs FP64; ensemble p50/p99 step latency"
# launch choices recorded by the autotuner on gfx950 for the bench workloads
TUNE_FILE = os.path.join(ROOT, "lesgo_amd", "tuned_gfx950.json")
TUNE_MODE = "auto"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3

WORKLOADS = {
    # name: (nx, ny, Nz, dt, seed, turbines)
    "windfarm256": (256, 256, 192, 1e-4, 7, True),
    "channel128": (128, 128, 128, 2e-4, 0, False),
    "channel32": (32, 32, 32, 2e-4, 0, False),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_case(name, precision):
    import lesgo_amd
    nx, ny, Nz, dt, seed, turb = WORKLOADS[name]
    cfg = lesgo_amd.Config(nx, ny, Nz, dt=dt, precision=precision, initu=False)
    c = lesgo_amd.Case(cfg)
    c.seed_state(seed)
    if turb:
        c.turbines_init(lesgo_amd.aligned_farm(cfg))
    return cfg, c


def measured_traffic(name, precision, kernel):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    PMC summary of this workload (FETCH_SIZE and WRITE_SIZE, calibrated on a
    plain copy), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{name}_{precision}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        f = ln.split()
        if len(f) >= 4 and f[0] == "traffic_per_launch_bytes" and f[1] == precision \
                and f[2] == kernel:
            return float(f[3])
    return None


def prepare_tuning(c):
    """auto: the recorded launch choices when the file exists, else tune this
    case now; tune: tune now; off: the built-in heuristics."""
    from lesgo_amd.case import tune_load
    if TUNE_MODE == "off":
        return "off"
    if TUNE_MODE == "auto" and os.path.exists(TUNE_FILE):
        tune_load(TUNE_FILE)
        return "file"
    c.autotune(5)
    return "live"


def override_variants(v):
    """Every recorded four-step launch choice with kernel variant v instead,
    where the library accepts it for that pass."""
    from lesgo_amd._lib import LesgoError
    from lesgo_amd.case import tune_set, tune_table
    for e in tune_table():
        if e["n1"] and e["variant"] != v:
            try:
                tune_set(e["kind"], e["precision"], e["N"], e["other"], e["n1"], e["n2"],
                         e["per_wg"], v)
            except LesgoError:
                pass


def step_figures(name, precision, ms_per_step):
    """Whole-step roofline: the algorithmic HBM bytes and FFT flops of one
    step (compulsory traffic with plane-local fusion and the one z barrier of
    the pressure tridiagonal: 31.8 full-field passes, FP64 254.4 B/pt,
    mixed 27.8 passes at 4 B + 4 at 8 B = 143.2 B/pt; flops
    N [54.8 x 2.5 log2(nx ny) + 9 x 2.25 x 2.5 log2(nx2 ny2)]) against the
    HBM and vector peaks."""
    import math
    nx, ny, Nz = WORKLOADS[name][:3]
    npts = nx * ny * Nz
    b_alg = npts * (254.4 if precision == "fp64" else 143.2)
    f_alg = npts * (54.8 * 2.5 * math.log2(nx * ny)
                    + 9 * 2.25 * 2.5 * math.log2(1.5 * nx * 1.5 * ny))
    peak_f = FP64_PEAK_TFLOPS if precision == "fp64" else FP32_PEAK_TFLOPS
    t = ms_per_step * 1e-3
    hbm = b_alg / (HBM_PEAK_GBS * 1e9 * t)
    flop = f_alg / (peak_f * 1e12 * t)
    return {"B_alg_bytes": b_alg, "hbm_frac": hbm, "F_alg_flop": f_alg,
            "flop_frac": flop, "peak_tflops": peak_f, "frac": max(hbm, flop)}


def timed_passes(name, precision, c, steps, warmup):
    """From the case's current state: warm-up, then three passes of `steps`
    steps each: graph-replayed with no per-launch events (the headline),
    eager with no events, and eager with HIP events around every launch of
    the transform and tridiagonal kernel classes (the per-kernel figures)."""
    nx, ny, Nz = WORKLOADS[name][:3]
    npts = nx * ny * Nz
    c.step_graph(warmup)
    c.synchronize()
    t0 = time.perf_counter()
    c.step_graph(steps)
    c.synchronize()
    t_graph = time.perf_counter() - t0
    t0 = time.perf_counter()
    c.step(steps)
    c.synchronize()
    t_eager = time.perf_counter() - t0
    c.timing(True)
    c.synchronize()
    t0 = time.perf_counter()
    c.step(steps)
    c.synchronize()
    t_events = time.perf_counter() - t0
    rep = c.timing_report()
    c.timing(False)
    rms = c.rmsdiv()
    ke = c.energy()
    if not (ke == ke and ke < 1e6):
        raise RuntimeError(f"{name} {precision}: the run diverged (energy {ke})")
    ms = 1e3 * t_graph / steps
    res = {
        "value": npts * steps / t_graph / 1e6,
        "ms_per_step": ms,
        "eager_value": npts * steps / t_eager / 1e6,
        "eager_ms_per_step": 1e3 * t_eager / steps,
        "events_ms_per_step": 1e3 * t_events / steps,
        "kernel_classes_ms_per_step": sum(v[0] for v in rep.values()) / steps,
        "rmsdiv": rms,
        "energy": ke,
        "kernels": {k: {"ms_total": v[0], "launches": v[1], "gbytes": v[2] / 1e9}
                    for k, v in rep.items()},
        "step_roofline": step_figures(name, precision, ms),
    }
    dom = max(rep.items(), key=lambda kv: kv[1][0])
    kms, nl, by = dom[1]
    achieved = by / nl / (kms / nl * 1e-3) / 1e9 if kms > 0 else 0.0
    res["roofline"] = {
        "kernel": dom[0],
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": measured_traffic(name, precision, dom[0]),
        "alg_bytes_per_launch": by / nl if nl else 0.0,
        "launches": nl,
        "avg_launch_us": 1e3 * kms / nl if nl else 0.0,
        "share_of_step": kms / (1e3 * t_events),
    }
    return res


def mixed_vs_fp64(c64, c32, Nz):
    """Relative L2 of the mixed case's velocity against the FP64 case's on
    the device (u, v on levels 1..Nz, w on 1..Nz+1, x < nx)."""
    return {f: c32.diff_l2(c64, f, 1, Nz if f != "w" else Nz + 1) for f in ("u", "v", "w")}


ENSEMBLE = dict(members=48, nx=96, ny=96, Nz=64, dt=2e-4, steps=200, dia=0.15, hub=0.1,
                ct_prime=1.33)


def ensemble_members(n, rng_seed=0):
    """Per-member (seed, z0, mean_p_force_x, disks): seeds 0..n-1, z0
    log-uniform in [1e-5, 1e-3], mean pressure gradient uniform in [0.5, 1.5],
    every other member with a 2x2 aligned disk array."""
    import numpy as np
    rng = np.random.default_rng(rng_seed)
    zo = 10.0 ** rng.uniform(-5.0, -3.0, n)
    mpf = rng.uniform(0.5, 1.5, n)
    return [(m, float(zo[m]), float(mpf[m]), m % 2 == 1) for m in range(n)]


def ensemble_case(precision, members):
    import lesgo_amd
    E = ENSEMBLE
    cfg = lesgo_amd.Config(E["nx"], E["ny"], E["Nz"], dt=E["dt"], precision=precision,
                           initu=False)
    c = lesgo_amd.Case(cfg, members=len(members))
    farm = lesgo_amd.aligned_farm(cfg, 2, 2, dia=E["dia"], hub=E["hub"], sx_d=7.0, sy_d=5.0,
                                  ct_prime=E["ct_prime"])
    for m, (seed, zo, mpf, turb) in enumerate(members):
        c.set_member(m, zo, mpf, 0.0)
        c.seed_state(seed, 0.1, member=m)
        if turb:
            c.turbines_init(farm, member=m)
    c.set_counters(0, 0)
    return cfg, c


def ensemble_leg(steps):
    """48 independent 96x96x64 half channels in one batch: mixed precision,
    autotuned launch choices, graph-replayed steps; per-step latency from
    host clocks around each synchronised step; per-member error against the
    same batch in plain FP64 (a batch member computes exactly what that
    case computes alone)."""
    import numpy as np
    E = ENSEMBLE
    members = ensemble_members(E["members"])
    npts = E["nx"] * E["ny"] * E["Nz"] * len(members)
    cfg, c = ensemble_case("mixed", members)
    t0 = time.perf_counter()
    tuning = prepare_tuning(c)
    t_tune = time.perf_counter() - t0
    c.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        c.step_graph(1)
        c.synchronize()
        lat.append(time.perf_counter() - a)
    dt = time.perf_counter() - t0
    mixed = {m: {n: c.download(n, member=m) for n in ("u", "v", "w")} for m in range(len(members))}
    c.close()
    log(f"[bench] ensemble mixed: {npts * steps / dt / 1e6:.1f} Mpts*steps/s")
    cfg64, c64 = ensemble_case("fp64", members)
    t0 = time.perf_counter()
    c64.step_graph(steps)
    c64.synchronize()
    dt64 = time.perf_counter() - t0
    X = np.s_[1:cfg.Nz + 1, :, :cfg.nx]
    errs = []
    for m in range(len(members)):
        e = []
        for n in ("u", "v", "w"):
            r = c64.download(n, member=m)[X]
            e.append(float(np.linalg.norm(mixed[m][n][X] - r) / np.linalg.norm(r)))
        errs.append(max(e))
    c64.close()
    lat_ms = np.array(lat) * 1e3
    return {
        "workload": (f"{len(members)} x {E['nx']}x{E['ny']}x{E['Nz']} half channels, seeds 0-47, "
                     "z0 log-uniform [1e-5,1e-3], mean_p_force U[0.5,1.5], odd members 2x2 ADM "
                     f"D={E['dia']} Ct'={E['ct_prime']}, dt={E['dt']}"),
        "steps": steps,
        "value": npts * steps / dt / 1e6,
        "unit": "Mpts*steps/s",
        "step_ms_p50": float(np.percentile(lat_ms, 50)),
        "step_ms_p99": float(np.percentile(lat_ms, 99)),
        "fp64_value": npts * steps / dt64 / 1e6,
        "tuning": tuning,
        "tuning_s": t_tune,
        "member_vel_rel_l2_vs_fp64": {"max": max(errs), "median": float(np.median(errs))},
    }


def serving_leg(minutes, saturation_jobs):
    """The serving loop on a Poisson stream (2 jobs/s, grids {64^3, 96x48x64,
    128x64x96, 128^3}, 200-2000 steps, z0 log-uniform, 30 % with a 2x2 disk
    array, mixed precision, 200 GB budget), drained; then the same arrival
    process served one job at a time in plain FP64 (first in first out,
    measured), and the first `saturation_jobs` jobs submitted at once (the
    sustained job rate)."""
    from lesgo_amd import serve
    jobs = serve.job_stream(minutes, 2.0, seed=0)
    res = serve.run_stream(jobs, 200.0, 16, progress=log)
    log(f"[bench] served: job p50 {res['job_latency_s_p50']:.2f} s; serial FP64 replay")
    ser = serve.run_serial_fp64(jobs, progress=log)
    res["fp64_serial"] = ser
    res["job_latency_p50_serial_over_served"] = ser["job_latency_s_p50"] / res["job_latency_s_p50"]
    res["job_latency_p99_serial_over_served"] = ser["job_latency_s_p99"] / res["job_latency_s_p99"]
    if saturation_jobs > 0:
        log(f"[bench] saturation: {saturation_jobs} jobs at once")
        res["saturation"] = serve.saturation(jobs[:saturation_jobs], 200.0, 16, progress=log)
    return res


CPU_SAMPLE_NZ = 48


def cpu_baseline(name):
    """The numpy oracle on this host's cores, timed over five consecutive
    steps of a bounded sample of the workload: the same horizontal grid,
    disks and level spacing with the lowest CPU_SAMPLE_NZ of its levels
    (Lz scaled with them, so dz is the workload's), run as a fresh start
    with DYN_init = 5 so that the five steps are four Smagorinsky steps and
    one Lagrangian dynamic step, the workload's 1-in-cs_count cadence.
    Every x-y plane costs the same, so Mpts*steps/s carries over.  Timed
    twice: on one thread (numpy's FFTs), and with the FFTs on every core
    this process may use (at most 16; scipy's threaded pocketfft, the rest
    of the step on one thread); `value` is the faster."""
    from oracle import lesgo_oracle as o
    nx, ny, Nz, dt, seed, turb = WORKLOADS[name]
    nzs = min(Nz, CPU_SAMPLE_NZ)
    ncores = min(16, len(os.sched_getaffinity(0)))

    def run(workers):
        o.set_fft_workers(workers)
        try:
            p = o.Params(nx, ny, nzs, lz=nzs / Nz, dt=dt, initu=False, dyn_init=5)
            sp = o.Spectral(p)
            s = o.initial_condition(p, seed)
            if turb:
                s.turbines = o.Turbines(p, o.aligned_farm(p))
            per = []
            for _ in range(5):
                a = time.perf_counter()
                o.step(sp, s)
                per.append(time.perf_counter() - a)
        finally:
            o.set_fft_workers(1)
        return nx * ny * nzs * 5 / sum(per) / 1e6, per

    v1, per1 = run(1)
    vn, pern = run(ncores) if ncores > 1 else (v1, per1)
    fmt = lambda per: ", ".join(f"{x:.1f}" for x in per)  # noqa: E731
    return {
        "value": max(v1, vn),
        "unit": "Mpts*steps/s",
        "cores": ncores if vn > v1 else 1,
        "kind": "port",
        "value_1core": v1,
        f"value_fft_on_{ncores}_cores": vn,
        "sample": (f"numpy FP64 oracle (pocketfft), 5 consecutive steps of {nx}x{ny}x{nzs}: the "
                   f"workload's grid, disks and dz on its lowest {nzs} levels (the 5th step runs "
                   f"the Lagrangian model); one thread: {fmt(per1)} s; FFTs on {ncores} threads "
                   f"(scipy.fft workers), the rest on one: {fmt(pern)} s"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="windfarm256", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--ensemble-steps", type=int, default=ENSEMBLE["steps"],
                    help="steps of the 48-member ensemble leg (0 skips it)")
    ap.add_argument("--serve-minutes", type=float, default=1.0,
                    help="length of the Poisson job stream of the serving leg (0 skips it)")
    ap.add_argument("--saturation-jobs", type=int, default=40,
                    help="jobs of the stream submitted at once for the serving loop's "
                         "sustained rate (0 skips it)")
    ap.add_argument("--tune", default="auto", choices=["auto", "tune", "off"],
                    help="launch choices: recorded file if present (auto), tune now, or "
                         "built-in heuristics")
    ap.add_argument("--write-tune-file", action="store_true",
                    help="only autotune the bench workloads and record the choices")
    ap.add_argument("--spinup", type=int, default=None,
                    help="untimed steps before the warm-up (default: DYN_init, so the timed "
                         "steps include the dynamic model)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    ap.add_argument("--variant-override", type=int, default=None, metavar="V",
                    help="after loading the launch choices, run every four-step pass with "
                         "kernel variant V where it is launchable (2: the unspecialised "
                         "kernels), for before/after profiles")
    args = ap.parse_args()
    if args.calibrate_copy:
        import ctypes
        from lesgo_amd import _lib
        ms = ctypes.c_double()
        nbytes = args.calibrate_copy * (1 << 20)
        _lib.check(_lib.lib().lg_calibrate_copy(nbytes, 5, ctypes.byref(ms)), "calibrate_copy")
        print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": ms.value,
                          "copy_GBps": 2 * nbytes / (ms.value * 1e-3) / 1e9}), flush=True)
        return
    global TUNE_MODE
    TUNE_MODE = args.tune
    if args.write_tune_file:
        from lesgo_amd.case import tune_clear, tune_save
        tune_clear()
        for prec in ("fp64", "mixed"):
            _, c = make_case("windfarm256", prec)
            c.autotune(7)
            c.close()
            _, c = ensemble_case(prec, ensemble_members(ENSEMBLE["members"]))
            c.autotune(7)
            c.close()
        tune_save(TUNE_FILE)
        print(json.dumps({"tune_file": TUNE_FILE}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs a case on one GPU, with no sharding; "
                                   "run with --gpus 1 (multi-GPU scaling is not built)"}))
        sys.exit(2)
    nx, ny, Nz, dt, seed, turb = WORKLOADS[args.workload]
    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    cases = {}
    for prec in precs:
        cfg, cases[prec] = make_case(args.workload, prec)
    tuning = prepare_tuning(cases[precs[0]])
    if args.variant_override is not None:
        tuning = f"{tuning}, variant {args.variant_override} forced"
        override_variants(args.variant_override)
    spin = cfg.dyn_init if args.spinup is None else args.spinup
    for prec in precs:
        log(f"[bench] {prec} {args.workload}: {spin} untimed spin-up steps")
        cases[prec].step_graph(spin)
        cases[prec].synchronize()
    accuracy = None
    if len(precs) == 2:
        accuracy = {"after_steps": spin,
                    "vel_rel_l2": mixed_vs_fp64(cases["fp64"], cases["mixed"], Nz)}
        log(f"[bench] mixed vs fp64 after {spin} steps: {accuracy['vel_rel_l2']}")
    out = {}
    for prec in precs:
        log(f"[bench] {prec}: warmup {args.warmup}, 3 x {args.steps} timed steps")
        out[prec] = timed_passes(args.workload, prec, cases[prec], args.steps, args.warmup)
        out[prec]["tuning"] = tuning
        cases[prec].close()
        log(f"[bench] {prec} done: {out[prec]['value']:.1f} Mpts*steps/s graph-replayed, "
            f"{out[prec]['eager_value']:.1f} eager")
    head = out.get("mixed") or out["fp64"]
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "Mpts*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if "mixed" in out else "fp64",
        "data": f"synthetic: seeded log-law profile + 10% uniform noise (seed {seed})",
        "config": {
            "workload": (f"{args.workload}: {nx}x{ny}x{Nz}, Lx=2pi Ly=pi Lz=1, dt={dt}, "
                         "Lagrangian scale-dependent SGS, equilibrium wall z0=1e-4"
                         + (", 4x4 ADM D=0.1 hub=0.1 7Dx5D Ct'=1.33" if turb else "")
                         + f"; fresh start (Euler step 1, Cs^2=0.03 until DYN_init="
                         f"{cfg.dyn_init}, then the dynamic model every 5 steps), {spin} "
                         f"untimed spin-up steps, {args.warmup} warm-up steps, timed steps "
                         "graph-replayed (lg_step_graph) with no per-launch events"),
            "grid": [nx, ny, Nz],
            "precision_of_value": "mixed" if "mixed" in out else "fp64",
        },
        "roofline": {k: v for k, v in head["roofline"].items()},
        "step_roofline": head["step_roofline"],
        "eager_value": head["eager_value"],
        "timing_ms_per_step": {k: head[k] for k in ("ms_per_step", "eager_ms_per_step",
                                                    "events_ms_per_step",
                                                    "kernel_classes_ms_per_step")},
    }
    if accuracy is not None:
        line["mixed_vs_fp64"] = accuracy
    if "fp64" in out:
        line["fp64"] = {k: out["fp64"][k] for k in ("value", "ms_per_step", "eager_value",
                                                    "eager_ms_per_step", "roofline",
                                                    "step_roofline")}
    if "fp64" in out and "mixed" in out:
        line["mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    line["kernels"] = head["kernels"]
    if args.ensemble_steps > 0:
        log(f"[bench] ensemble: {ENSEMBLE['members']} members, {args.ensemble_steps} steps")
        line["ensemble"] = ensemble_leg(args.ensemble_steps)
        log(f"[bench] ensemble done: {line['ensemble']['value']:.1f} Mpts*steps/s, step p50 "
            f"{line['ensemble']['step_ms_p50']:.2f} ms p99 {line['ensemble']['step_ms_p99']:.2f} ms")
    if args.serve_minutes > 0:
        log(f"[bench] serving: {args.serve_minutes} min Poisson stream at 2 jobs/s")
        line["serving"] = serving_leg(args.serve_minutes, args.saturation_jobs)
        log(f"[bench] serving done: job p50 {line['serving']['job_latency_s_p50']:.2f} s "
            f"p99 {line['serving']['job_latency_s_p99']:.2f} s")
    if not args.no_cpu:
        log(f"[bench] cpu baseline (numpy oracle, {CPU_SAMPLE_NZ} levels, 5 steps)")
        line["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
