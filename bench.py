#!/usr/bin/env python3
"""Throughput of the SPH cutting time step on one MI355X, in particle-steps/s.

Default workload: the 1,000,000-particle single-grain cut (200 x 100 x 50 lattice, dx = 5 um,
rounded truncated-octahedron grain of ~4,800 triangles, PSE heat with tool-workpiece thermal
contact, CSPM, Gray artificial stress, Johnson-Cook plasticity), FP64 method.  Synthetic inputs
are built by the library's set-up code (seeded); everything the timed steps touch is resident in
HBM before timing starts.

One JSON line on stdout: metric, value (particle-steps/s of the 1M FP64 run), per-kernel
timings from HIP events on the library stream over the timed region, the rooflines of the
dominant kernel (the pair pass) and of the neighbour pass, the whole step's VALU and HBM
fractions, the metric's FP32 + Kahan vs FP64 pairs at 1M (same workload) and 8M particles (three
grains), the p50 / p99 time to result of a short served stream, and the CPU oracle timed on the
host cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks
HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6
FP32_VECTOR_PEAK_TFLOPS = 157.3

# algorithmic work per unit:
#   pair pass: 280 FP operations per interacting pair (CSPM + B 55, grad W 25, grad v 30,
#              artificial stress 15, stress divergence 63, viscosity 40, XSPH 10, Brookshaw 25)
#   neighbour pass: 11 per candidate (r^2 and support test) + 10 per candidate for the PSE sum
#   algorithmic HBM bytes of the pair pass per particle: its own record read once (192 B FP64,
#   112 B FP32; every neighbour record is a record some particle reads anyway), its 32-byte
#   CSPM entry read once, flags and list length (8 B), its 25 FP64 neighbour sums written
#   (200 B), and its list (4 B per pair)
FLOPS_PER_PAIR = 280.0
FLOPS_PER_CANDIDATE = 11.0
FLOPS_PER_PSE_CANDIDATE = 10.0


WORKLOADS = {
    (1, 1.0): "orthogonal cut, 8,000-particle block (40x20x10, dx 10 um), 12-triangle wedge, Brookshaw heat",
    (2, 1.0): "orthogonal cut, 1M-particle block (200x100x50, dx 5 um), single synthetic grain "
              "(4,800 triangles), PSE heat with tool-workpiece thermal contact",
    (3, 1.0): "orthogonal cut, 8M-particle block (400x200x100, dx 2.5 um), three synthetic grains",
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def gpu_run(args, preset=None, scale=None, precision=None, steps=None, warmup=None, scene=None):
    """Times `steps` steps of one workload after `warmup` untimed ones. HIP events on the library
    stream around the neighbour and pair passes time those kernels over the timed region; the
    breakdown of every kernel comes from BREAKDOWN_STEPS further, untimed steps with events around
    every launch (such events add 0.05-0.1 ms to a 1M-particle step, so the timed region carries
    only the two passes the roofline figures are quoted on)."""
    from iwf_mfree_gpu_3d_amd import _lib, api, scenes
    _lib.init(0)
    preset = args.preset if preset is None else preset
    scale = args.scale if scale is None else scale
    precision = args.precision if precision is None else precision
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    t0 = time.time()
    if scene is None:
        scene = scenes.Scene(preset, scale, args.seed)
    log(f"scene preset {preset} scale {scale}: {scene.n} particles "
        f"({scene.params.n_workpiece} workpiece), {sum(t[1].shape[0] for t in scene.tools)} tool triangles, "
        f"built in {time.time() - t0:.1f}s")
    prec = {"fp64": _lib.MFG_FP64, "fp32-kahan": _lib.MFG_FP32_KAHAN, "fp32-wide": _lib.MFG_FP32_WIDE}[precision]
    parts, grid = scenes.setup_cut(scene, precision=prec, single_sqrt=True, cell_tile=args.cell_tile,
                                   sort_interval=args.sort_interval)
    lf = api.leap_frog(scene.n, grid.num_cell())
    api.global_time_current = api.global_time_dt
    scenes.run(lf, parts, grid, warmup)
    api.synchronize()
    n_before = parts.N
    api.profile_reset()
    api.profile_enable(_lib.MFG_PROFILE_PASSES)
    api.synchronize()
    t0 = time.perf_counter()
    scenes.run(lf, parts, grid, steps - 1)
    api.pair_stats_enable(True)  # the last timed step counts its pairs (the roofline's work)
    scenes.run(lf, parts, grid, 1)
    api.synchronize()
    t1 = time.perf_counter()
    api.pair_stats_enable(False)
    api.profile_enable(False)
    prof = api.profile_read()
    n_after = parts.N
    cand, pairs = parts.pair_stats()
    api.profile_reset()
    api.profile_enable(_lib.MFG_PROFILE_ALL)
    scenes.run(lf, parts, grid, min(steps, BREAKDOWN_STEPS))
    api.synchronize()
    api.profile_enable(False)
    brk = api.profile_read()
    secs = t1 - t0
    n_active = 0.5 * (n_before + n_after)
    value = n_active * steps / secs
    kernels = {k: {"launches": v[0], "avg_ms": v[1] / max(v[0], 1), "total_ms": v[1]} for k, v in prof.items()}
    breakdown = {k: {"launches": v[0], "avg_ms": v[1] / max(v[0], 1), "total_ms": v[1]} for k, v in brk.items()}
    log(f"  preset {preset} {precision}: {value:.4e} particle-steps/s ({secs / steps * 1e3:.3f} ms/step)")
    del parts, grid, lf
    return dict(scene=scene, value=value, secs=secs, steps=steps, n_active=n_active, kernels=kernels,
                breakdown=breakdown, cand=cand, pairs=pairs, n_particles=scene.n, precision=precision, preset=preset, scale=scale)


# untimed steps after the timed region with events around every kernel (the per-kernel breakdown)
BREAKDOWN_STEPS = 20

# algorithmic HBM bytes of one pair pass per particle, besides 4 B per neighbour-list entry:
#   own pair record (192 B FP64 / 112 B FP32) + CSPM entry 32 B + count 4 B + flags 4 B + 25
#   neighbour sums written (200 B); the particle update that reads the sums back is a kernel of
#   its own
PAIR_BYTES_PER_PARTICLE = {"fp64": 432.0, "fp32": 352.0}

# committed rocprofv3 FETCH_SIZE / WRITE_SIZE summaries: (kernel, preset, scale) -> file
TRAFFIC_PROFILES = {
    ("pair_fp64", 2, 1.0): "pair_fp64_cut1M_single_grain_traffic.md",
    ("pair_fp32_kahan", 2, 1.0): "pair_fp32_kahan_cut1M_single_grain_traffic.md",
}


def measured_traffic(kernel: str, args):
    """HBM-side bytes per launch of `kernel` from the committed profiler summary of this workload
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of known size), or None."""
    name = TRAFFIC_PROFILES.get((kernel, args.preset, float(args.scale)))
    if name is None:
        return None
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch: (\d+)", open(path).read())
    return int(m.group(1)) if m else None


# algorithmic bytes per particle-step of the whole step (two-phase step, every field once per
# phase, plus the re-sort every step): FP64 1,062 + 694 B; FP32 pair arithmetic with FP64
# positions 594 + 398 B
STEP_BYTES = {"fp64": 1062.0 + 694.0, "fp32": 594.0 + 398.0}
# FP operations per particle-step outside the pair and neighbour passes: predict, EOS, the 3x3
# eigen solve and artificial stress, Jaumann rate, the update and the Johnson-Cook return
FLOPS_PER_PARTICLE_UPDATE = 400.0


def kernel_roofline(res, kernel: str, args, pse: bool):
    """Roofline of one kernel of the step: the pair and neighbour passes are bound by the vector
    ALU (no step of the method is a dense contraction: no matrix instructions), with the HBM
    view (algorithmic bytes over the launch time, and the measured traffic) beside it."""
    kernels = res["kernels"]
    if kernel not in kernels:
        return None
    avg_s = kernels[kernel]["avg_ms"] * 1e-3
    fp64 = res["precision"] == "fp64"
    peak = FP64_VECTOR_PEAK_TFLOPS if fp64 else FP32_VECTOR_PEAK_TFLOPS
    n = res["n_active"]
    if kernel.startswith("pair_"):
        flops = FLOPS_PER_PAIR * res["pairs"]
        basis = f"{FLOPS_PER_PAIR:.0f} flop x {res['pairs']} interacting pairs per launch"
        alg_bytes = PAIR_BYTES_PER_PARTICLE["fp64" if fp64 else "fp32"] * n + 4.0 * res["pairs"]
    elif kernel == "neighbours":
        per = FLOPS_PER_CANDIDATE + (FLOPS_PER_PSE_CANDIDATE if pse else 0.0)
        flops = per * res["cand"]
        basis = f"{per:.0f} flop x {res['cand']} candidate pairs per launch"
        alg_bytes = 76.0 * n + 4.0 * res["pairs"]
    else:
        return None
    achieved = flops / avg_s / 1e12
    hbm = alg_bytes / avg_s / 1e9
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": measured_traffic(kernel, args), "kernel": kernel,
            "basis": basis, "avg_launch_ms": kernels[kernel]["avg_ms"],
            "compute_units": ("FP64 vector ALU (78.6 TFLOP/s published; the FP64 matrix peak is the same "
                              "figure and no matrix instructions are used)" if fp64 else
                              "FP32 vector ALU, packed FP32 not generated; peak quoted is the published "
                              "157.3 TFLOP/s FP32 figure"),
            "algorithmic_bytes": alg_bytes, "hbm_achieved_GBps": hbm, "hbm_peak_GBps": HBM_PEAK_GBS,
            "hbm_frac": hbm / HBM_PEAK_GBS}


def whole_step(res, pse: bool):
    """VALU and HBM fractions of the whole step: algorithmic flops (pair terms, candidate tests and
    PSE terms, per-particle update) and algorithmic bytes per particle-step over the measured
    time per step."""
    fp64 = res["precision"] == "fp64"
    n = res["n_active"]
    step_s = res["secs"] / res["steps"]
    flops = (FLOPS_PER_PAIR * res["pairs"] + (FLOPS_PER_CANDIDATE + (FLOPS_PER_PSE_CANDIDATE if pse else 0.0))
             * res["cand"] + FLOPS_PER_PARTICLE_UPDATE * n)
    bytes_ = STEP_BYTES["fp64" if fp64 else "fp32"] * n
    peak = FP64_VECTOR_PEAK_TFLOPS if fp64 else FP32_VECTOR_PEAK_TFLOPS
    return {"ms_per_step": step_s * 1e3, "valu_TFLOPs": flops / step_s / 1e12, "valu_frac": flops / step_s / 1e12 / peak,
            "hbm_GBps": bytes_ / step_s / 1e9, "hbm_frac": bytes_ / step_s / 1e9 / HBM_PEAK_GBS,
            "flops_per_particle_step": flops / n, "bytes_per_particle_step": STEP_BYTES["fp64" if fp64 else "fp32"]}


def cpu_baseline(args):
    """The FP64 oracle (a CPU restatement of the reference's step) built with g++ -O3
    -march=native -fopenmp for this host, timed on its cores on a bounded sample of the same
    workload: all cores (OMP_NUM_THREADS, else the CPU count, OMP_PROC_BIND=close) for
    --cpu-steps steps, and one thread for one step.  Each run is a process of its own, so the
    OpenMP runtime starts with its thread settings."""
    import subprocess
    import oracle
    lib_path = oracle.build_native()
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        m = re.search(r"Model name:\s*(.+)", out)
        model = m.group(1).strip() if m else model
    except (OSError, subprocess.SubprocessError):
        pass
    nproc = os.cpu_count() or 1
    all_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc

    def run(threads, steps):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", MFG_ORACLE_LIB=lib_path)
        r = subprocess.run([sys.executable, "-m", "oracle.cpu_bench", "--preset", str(args.preset), "--scale",
                            str(args.scale), "--seed", str(args.seed), "--steps", str(steps), "--threads",
                            str(threads)], capture_output=True, text=True, env=env, cwd=ROOT, timeout=900)
        if r.returncode != 0:
            raise RuntimeError(f"CPU baseline run failed: {r.stderr[-2000:]}")
        return json.loads(r.stdout.strip().splitlines()[-1])

    many = run(all_threads, args.cpu_steps)
    one = run(1, 1) if args.cpu_single_steps > 0 else None
    res = {"value": many["value"], "unit": "particle-steps/s", "cores": all_threads, "kind": "port",
           "sample": (f"same workload ({many['particles']} particles): {args.cpu_steps} steps after 1 warm-up "
                      f"step on {all_threads} threads ({many['seconds']:.1f} s)"
                      + (f"; 1 step on 1 thread ({one['seconds']:.1f} s)" if one else "")),
           "build": "g++ -O3 -march=native -fopenmp -ffp-contract=off", "omp_proc_bind": "close",
           "nproc": nproc, "cpu_model": model}
    if one:
        res["value_1_thread"] = one["value"]
    return res


def copy_probe(mib: int) -> None:
    """A plain device copy of known size (y = 2 x over `mib` MiB of float32), used to scale the
    FETCH_SIZE / WRITE_SIZE counters of the profiler to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.rand(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 5}))


def error_report(args) -> None:
    """The precision form: the FP32 modes against the FP64 run of the same workload after
    --steps steps, particles matched by persistent id; relative L2 error
    ||x - x_ref|| / ||x_ref|| and max absolute error of stress, plastic strain, temperature rise
    and position."""
    import numpy as np
    from iwf_mfree_gpu_3d_amd import _lib, api, scenes
    _lib.init(0)
    fields = {"S": _lib.F_S, "eps_pl": _lib.F_EPS_PL, "T": _lib.F_T, "pos": _lib.F_POS}
    runs, times = {}, {}
    for name, prec in [("fp64", _lib.MFG_FP64), ("fp32-kahan", _lib.MFG_FP32_KAHAN),
                       ("fp32-wide", _lib.MFG_FP32_WIDE)]:
        scene = scenes.Scene(args.preset, args.scale, args.seed)
        parts, grid = scenes.setup_cut(scene, precision=prec, single_sqrt=True)
        lf = api.leap_frog(scene.n, grid.num_cell())
        api.global_time_current = api.global_time_dt
        api.synchronize()
        t0 = time.perf_counter()
        scenes.run(lf, parts, grid, args.steps)
        api.synchronize()
        times[name] = time.perf_counter() - t0
        order = np.argsort(parts.download(_lib.F_ID).astype(np.int64))
        runs[name] = {k: parts.download(f)[order] for k, f in fields.items()}
        runs[name]["T"] = runs[name]["T"] - scene.params.trml_wp.T_init
        log(f"{name}: {scene.n} particles, {args.steps} steps in {times[name]:.2f} s")
        del parts, grid, lf
    ref = runs["fp64"]
    errors = {}
    for name in ("fp32-kahan", "fp32-wide"):
        e = {}
        for k in fields:
            a, b = runs[name][k], ref[k]
            nb = np.linalg.norm(b.ravel())
            e[k] = {"l2_rel": float(np.linalg.norm((a - b).ravel()) / nb) if nb > 0 else None,
                    "max_abs": float(np.abs(a - b).max())}
        errors[name] = e
        log(f"{name} vs fp64: " + ", ".join(f"{k} L2 {v['l2_rel']:.2e} max {v['max_abs']:.2e}" for k, v in e.items()))
    n = scene.n
    out = {
        "metric": "particle-steps/s at 1M/8M particles, FP32 vs FP64; p99 sweep latency",
        "value": n * args.steps / times["fp32-kahan"], "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 0, "ms_per_step": times["fp32-kahan"] / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded lattice + jitter, synthetic grain meshes)",
        "config": {"workload": WORKLOADS.get((args.preset, float(args.scale)), f"preset {args.preset}"),
                   "particles": n, "precision": "fp32-kahan vs fp64"},
        "particle_steps_per_s": {k: n * args.steps / v for k, v in times.items()},
        "errors_vs_fp64": errors,
        "error_note": "T is the temperature rise T - T_init; L2 relative = ||x - x_fp64|| / ||x_fp64||",
    }
    print(json.dumps(out))


def sweep(args) -> None:
    """Served or batched form: one JSON line with the aggregate particle-steps/s as the value and
    the p50 / p99 time to result beside it."""
    from iwf_mfree_gpu_3d_amd import _lib, serve
    _lib.init(0)
    if args.batched > 0:
        reqs = serve.sweep_requests(count=args.batched, batched=True, seed=args.seed, n_min=5e3, n_max=2e5,
                                    steps=(300, 300), fp32_share=1.0)
        workload = (f"batched: {args.batched} orthogonal-cut instances at t=0, N log-uniform in [5e3, 2e5] "
                    f"(4:2:1 blocks), depth U[5,40] um, speed U[5,50] m/s, rake U[-30,10] deg, JC A/B/C x U(0.9,1.1), "
                    f"FP32+Kahan, 300 steps each")
    else:
        reqs = serve.sweep_requests(rate=4.0, duration=args.served, seed=args.seed)
        workload = (f"served: Poisson arrivals 4/s over {args.served:g} s, N log-uniform in [2e4, 2e6] (4:2:1 blocks), "
                    f"70% FP32+Kahan / 30% FP64, 100-1000 steps, parameters as the batched form")
    log(f"{len(reqs)} requests, {sum(r.lattice()[0] * r.lattice()[1] * r.lattice()[2] for r in reqs)} particles in all")
    mode = args.mode
    tuning, tune_s = None, None
    if args.autotune:
        from iwf_mfree_gpu_3d_amd import autotune
        t0 = time.perf_counter()
        # before the timed request loop
        tuning = autotune.tune(reqs, log=log) if mode == "streams" else autotune.tune_batched(reqs, log=log)
        tune_s = time.perf_counter() - t0
        log(f"autotune: {tune_s:.1f} s")
    res = serve.serve(reqs, hbm_budget=args.hbm_budget_gb * 1e9, log=log, streams=args.streams,
                      tuning=tuning, mode=mode, use_graph=not args.no_graph, cell_tile=args.cell_tile)
    log(f"p50 {res['p50_s']:.2f} s, p99 {res['p99_s']:.2f} s, {res['particle_steps_per_s']:.4e} particle-steps/s "
        f"over {res['makespan_s']:.1f} s")
    out = {
        "metric": "particle-steps/s at 1M/8M particles, FP32 vs FP64; p99 sweep latency",
        "value": res["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1,
        "steps": int(sum(r.steps for r in reqs)), "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "mixed" if args.served > 0 else "f32",
        "data": "synthetic (seeded lattices + jitter, wedge tools)",
        "config": {"workload": workload, "requests": len(reqs), "hbm_budget_gb": args.hbm_budget_gb,
                   "streams": args.streams, "autotune": bool(args.autotune), "mode": mode,
                   "graph": mode == "batched" and not args.no_graph, "cell_tile": args.cell_tile},
        "sweep": {"p50_time_to_result_s": res["p50_s"], "p99_time_to_result_s": res["p99_s"],
                  "makespan_s": res["makespan_s"], "particle_steps": res["particle_steps"],
                  "setup_s": res["setup_s"],
                  "particle_steps_per_s_excluding_setup": res["particle_steps"] / max(res["makespan_s"] - res["setup_s"], 1e-9)},
    }
    if tuning is not None:
        out["autotune"] = {"seconds": tune_s,
                           "choices": [{"bucket_log2_particles": b, "precision": p, "pair_pass": c[0],
                                        "strip": c[1], "cell_tile": c[2], "sort_interval": max(c[3], 1),
                                        "pair_block": c[4] or 256}
                                       for (b, p), c in sorted(tuning.choice.items())],
                           "batch_cell_tile": {str(p): t for p, t in sorted(tuning.batch_tile.items())},
                           "batch_ms_per_step": {f"{p}/{t}": s * 1e3 for (p, t), s in sorted(tuning.batch_times.items())}}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--preset", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "fp32-kahan", "fp32-wide"], default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=8,
                    help="steps of the CPU oracle timed on all cores for cpu_baseline (about 1 s each at "
                         "1M particles on 16 threads)")
    ap.add_argument("--cpu-single-steps", type=int, default=1,
                    help="steps of the CPU oracle timed on one thread (0: skip)")
    ap.add_argument("--steps-8m", type=int, default=10,
                    help="timed steps of the 8M-particle three-grain runs (FP64 and FP32 + Kahan)")
    ap.add_argument("--warmup-8m", type=int, default=3)
    ap.add_argument("--sweep-seconds", type=float, default=4.0,
                    help="seconds of Poisson arrivals of the served stream whose p50 / p99 time to result "
                         "the default line reports (0: skip)")
    ap.add_argument("--quick", action="store_true",
                    help="only the headline workload (no FP32 / 8M runs, no served stream)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain copy of this many MiB (profiler counter calibration)")
    ap.add_argument("--served", type=float, default=0.0,
                    help="served form instead: Poisson arrivals (4 requests/s) over this many seconds of "
                         "cutting-sweep requests, N log-uniform in [2e4, 2e6], 70%% FP32+Kahan, 100-1000 steps")
    ap.add_argument("--batched", type=int, default=0,
                    help="batched form instead: this many instances at time zero, N log-uniform in "
                         "[5e3, 2e5], FP32+Kahan, 300 steps")
    ap.add_argument("--hbm-budget-gb", type=float, default=250.0)
    ap.add_argument("--autotune", action="store_true",
                    help="served / batched forms: pick the pair-pass form and visiting order per size "
                         "bucket by measurement first (outside the timed loop)")
    ap.add_argument("--streams", type=int, default=4,
                    help="served / batched forms: streams the instances are spread over")
    ap.add_argument("--mode", choices=["batched", "streams"], default="batched",
                    help="served / batched forms: one launch per kernel over the instances of a precision "
                         "(batched) or each instance's own kernels on a stream pool (streams)")
    ap.add_argument("--cell-tile", type=int, default=0, choices=[0, 1, 2, 4, 8, 16],
                    help="cells per wavefront of the neighbour pass (0 automatic)")
    ap.add_argument("--sort-interval", type=int, default=0,
                    help="opt-in: sort every k-th step of the single-instance runs (0 or 1: every step, "
                         "the reference's schedule; other values change the results)")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched mode: launch every step's kernels instead of replaying a captured step")
    ap.add_argument("--batched-instances", type=int, default=256,
                    help="instances of the batched record in the default line (0: skip)")
    ap.add_argument("--error-report", action="store_true",
                    help="precision form instead: run the workload --steps steps in FP64 and in both FP32 "
                         "modes and report the FP32 L2 and max errors against the FP64 run")
    args = ap.parse_args()
    if args.steps < 1 or args.warmup < 0:
        ap.error("--steps must be at least 1 and --warmup at least 0")
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return
    if args.served > 0 or args.batched > 0:
        sweep(args)
        return
    if args.error_report:
        error_report(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(one process drives one GPU; the step would shard by slabs along the cut direction)", file=sys.stderr)
        sys.exit(2)
    res = gpu_run(args)
    scene = res["scene"]
    pse = scene.params.heat == 2
    dom = max(res["kernels"], key=lambda k: res["kernels"][k]["total_ms"])
    roof = kernel_roofline(res, dom, args, pse)
    nb_roof = kernel_roofline(res, "neighbours", args, pse)
    step = whole_step(res, pse)
    for k, v in sorted(res["kernels"].items(), key=lambda kv: -kv[1]["total_ms"]):
        log(f"  {k:20s} {v['launches']:6d} launches  avg {v['avg_ms']:.4f} ms  total {v['total_ms']:.1f} ms")
    log(f"  every kernel, {min(args.steps, BREAKDOWN_STEPS)} further steps with events around each launch:")
    for k, v in sorted(res["breakdown"].items(), key=lambda kv: -kv[1]["total_ms"]):
        log(f"    {k:20s} {v['launches']:6d} launches  avg {v['avg_ms']:.4f} ms")
    log(f"GPU: {res['value']:.4e} particle-steps/s over {args.steps} steps ({res['secs']:.3f} s)")
    precision_pairs, sweep_rec, batched_rec = None, None, None
    if not args.quick and args.preset == 2 and args.precision == "fp64":
        from iwf_mfree_gpu_3d_amd import scenes
        # the metric's FP32 vs FP64 pairs: the same 1M workload in FP32 + Kahan, and the 8M
        # three-grain workload in both modes
        r32 = gpu_run(args, precision="fp32-kahan", scene=scene)
        scene8 = scenes.Scene(3, 1.0, args.seed)
        r8 = {p: gpu_run(args, preset=3, scale=1.0, precision=p, steps=args.steps_8m, warmup=args.warmup_8m,
                         scene=scene8) for p in ("fp64", "fp32-kahan")}
        precision_pairs = {
            "1M": {"particles": res["n_particles"], "fp64": res["value"], "fp32_kahan": r32["value"],
                   "fp32_over_fp64": r32["value"] / res["value"], "steps": args.steps,
                   "fp32_pair_kernel_ms": r32["kernels"].get("pair_fp32_kahan", {}).get("avg_ms")},
            "8M": {"particles": scene8.n, "fp64": r8["fp64"]["value"], "fp32_kahan": r8["fp32-kahan"]["value"],
                   "fp32_over_fp64": r8["fp32-kahan"]["value"] / r8["fp64"]["value"], "steps": args.steps_8m,
                   "workload": WORKLOADS[(3, 1.0)],
                   "whole_step_fp64": whole_step(r8["fp64"], True),
                   "whole_step_fp32_kahan": whole_step(r8["fp32-kahan"], True)},
        }
        del scene8
        if args.sweep_seconds > 0:
            from iwf_mfree_gpu_3d_amd import serve
            reqs = serve.sweep_requests(rate=4.0, duration=args.sweep_seconds, seed=args.seed)
            t0 = time.perf_counter()
            sres = serve.serve(reqs, hbm_budget=args.hbm_budget_gb * 1e9, streams=args.streams)
            log(f"served stream: {len(reqs)} requests in {time.perf_counter() - t0:.1f} s, p50 {sres['p50_s']:.2f} s, "
                f"p99 {sres['p99_s']:.2f} s, {sres['particle_steps_per_s']:.4e} particle-steps/s")
            sweep_rec = {"workload": (f"served: Poisson arrivals 4/s over {args.sweep_seconds:g} s, N log-uniform in "
                                      f"[2e4, 2e6] (4:2:1 blocks), 70% FP32+Kahan / 30% FP64, 100-1000 steps"),
                         "requests": len(reqs), "p50_time_to_result_s": sres["p50_s"],
                         "p99_time_to_result_s": sres["p99_s"], "particle_steps_per_s": sres["particle_steps_per_s"],
                         "makespan_s": sres["makespan_s"], "setup_s": sres["setup_s"],
                         "particle_steps_per_s_excluding_setup":
                             sres["particle_steps"] / max(sres["makespan_s"] - sres["setup_s"], 1e-9),
                         "mode": "batched, graph-replayed steps"}
        if args.batched_instances > 0:
            from iwf_mfree_gpu_3d_amd import serve
            reqs = serve.sweep_requests(count=args.batched_instances, batched=True, seed=args.seed, n_min=5e3,
                                        n_max=2e5, steps=(300, 300), fp32_share=1.0)
            bres = serve.serve(reqs, hbm_budget=args.hbm_budget_gb * 1e9, streams=args.streams)
            log(f"batched: {len(reqs)} instances, {bres['particle_steps_per_s']:.4e} particle-steps/s, "
                f"p99 {bres['p99_s']:.2f} s over {bres['makespan_s']:.1f} s")
            batched_rec = {"workload": (f"batched: {len(reqs)} orthogonal-cut instances at t=0, N log-uniform in "
                                        "[5e3, 2e5], FP32+Kahan, 300 steps each; one launch per kernel over all "
                                        "instances, steps replayed from a captured graph"),
                           "particles": int(sum(r.lattice()[0] * r.lattice()[1] * r.lattice()[2] for r in reqs)),
                           "particle_steps_per_s": bres["particle_steps_per_s"], "p50_time_to_result_s": bres["p50_s"],
                           "p99_time_to_result_s": bres["p99_s"], "makespan_s": bres["makespan_s"],
                           "setup_s": bres["setup_s"],
                           "particle_steps_per_s_excluding_setup":
                               bres["particle_steps"] / max(bres["makespan_s"] - bres["setup_s"], 1e-9)}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args)
        log(f"CPU oracle: {cpu['value']:.4e} particle-steps/s on {cpu['cores']} threads")
    out = {
        "metric": "particle-steps/s at 1M/8M particles, FP32 vs FP64; p99 sweep latency",
        "value": res["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["secs"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded lattice + jitter, synthetic grain mesh)",
        "config": {"workload": WORKLOADS.get((args.preset, float(args.scale)),
                                             f"orthogonal cut preset {args.preset} (scale {args.scale})"),
                   "particles": res["n_particles"], "active_particles": res["n_active"],
                   "precision": args.precision, "seed": args.seed,
                   "sort_interval": max(args.sort_interval, 1),
                   "pairs_per_particle": res["pairs"] / max(res["n_active"], 1),
                   "candidates_per_particle": res["cand"] / max(res["n_active"], 1)},
        "roofline": roof,
        "roofline_neighbours": nb_roof,
        "whole_step": step,
        "precision_pairs": precision_pairs,
        "sweep": sweep_rec,
        "batched": batched_rec,
        "kernels": res["kernels"],
        "kernel_breakdown": {"steps": min(args.steps, BREAKDOWN_STEPS),
                             "note": "untimed steps after the timed region, events around every kernel",
                             "kernels": res["breakdown"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
