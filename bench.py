#!/usr/bin/env python3
"""Throughput of the SPH cutting time step on one MI355X, in particle-steps/s.

Default workload: the 1,000,000-particle single-grain cut (200 x 100 x 50 lattice, dx = 5 um,
rounded truncated-octahedron grain of ~4,800 triangles, PSE heat with tool-workpiece thermal
contact, CSPM, Gray artificial stress, Johnson-Cook plasticity), FP64 method.  Synthetic inputs
are built by the library's set-up code (seeded); everything the timed steps touch is resident in
HBM before timing starts.

One JSON line on stdout: metric, value (particle-steps/s), per-kernel timings from HIP events on
the library stream over the timed region, the roofline of the dominant kernel, and the CPU
oracle timed on the host cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
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
#   algorithmic HBM bytes of the pair pass per particle: its own record and state read once,
#   its neighbour list (4 B per pair), the updated state written once (FP64 field counts below)
FLOPS_PER_PAIR = 280.0
FLOPS_PER_CANDIDATE = 11.0
FLOPS_PER_PSE_CANDIDATE = 10.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def gpu_run(args):
    from iwf_mfree_gpu_3d_amd import _lib, api, scenes
    _lib.init(0)
    t0 = time.time()
    scene = scenes.Scene(args.preset, args.scale, args.seed)
    log(f"scene preset {args.preset} scale {args.scale}: {scene.n} particles "
        f"({scene.params.n_workpiece} workpiece), {sum(t[1].shape[0] for t in scene.tools)} tool triangles, "
        f"built in {time.time() - t0:.1f}s")
    prec = {"fp64": _lib.MFG_FP64, "fp32-kahan": _lib.MFG_FP32_KAHAN, "fp32-wide": _lib.MFG_FP32_WIDE}[args.precision]
    parts, grid = scenes.setup_cut(scene, precision=prec, single_sqrt=True)
    lf = api.leap_frog(scene.n, grid.num_cell())
    api.global_time_current = api.global_time_dt
    scenes.run(lf, parts, grid, args.warmup)
    api.synchronize()
    n_before = parts.N
    api.profile_reset()
    api.profile_enable(True)
    api.synchronize()
    t0 = time.perf_counter()
    scenes.run(lf, parts, grid, args.steps)
    api.synchronize()
    t1 = time.perf_counter()
    api.profile_enable(False)
    prof = api.profile_read()
    n_after = parts.N
    cand, pairs = parts.pair_stats()
    secs = t1 - t0
    n_active = 0.5 * (n_before + n_after)
    value = n_active * args.steps / secs
    kernels = {k: {"launches": v[0], "avg_ms": v[1] / max(v[0], 1), "total_ms": v[1]} for k, v in prof.items()}
    return dict(scene=scene, value=value, secs=secs, n_active=n_active, kernels=kernels, cand=cand, pairs=pairs,
                parts=parts, n_particles=scene.n)


def roofline(res, precision: str, pse: bool):
    kernels = res["kernels"]
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    avg_s = kernels[dom]["avg_ms"] * 1e-3
    peak = FP64_VECTOR_PEAK_TFLOPS if precision == "fp64" else FP32_VECTOR_PEAK_TFLOPS
    if dom.startswith("pair_update"):
        flops = FLOPS_PER_PAIR * res["pairs"]
        basis = f"{FLOPS_PER_PAIR:.0f} flop x {res['pairs']} interacting pairs per launch"
    elif dom == "neighbours":
        per = FLOPS_PER_CANDIDATE + (FLOPS_PER_PSE_CANDIDATE if pse else 0.0)
        flops = per * res["cand"]
        basis = f"{per:.0f} flop x {res['cand']} candidate pairs per launch"
    else:
        return dom, None
    achieved = flops / avg_s / 1e12
    return dom, {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                 "frac": achieved / peak, "traffic": None, "kernel": dom, "basis": basis,
                 "avg_launch_ms": kernels[dom]["avg_ms"]}


def cpu_baseline(args):
    """The FP64 oracle on the host cores: the same workload, a bounded number of steps."""
    import oracle
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import helpers
    from iwf_mfree_gpu_3d_amd import scenes
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    scene = scenes.Scene(args.preset, args.scale, args.seed)
    sim = helpers.oracle_from_scene(scene, single_sqrt=True, threads=threads)
    dt = scene.params.dt
    sim.set_time(dt, dt)
    sim.run(1)  # warm-up step (first sort from the unsorted lattice)
    t0 = time.perf_counter()
    sim.run(args.cpu_steps)
    secs = time.perf_counter() - t0
    n = sim.counts()[0]
    return {"value": n * args.cpu_steps / secs, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"same workload ({scene.n} particles), {args.cpu_steps} steps after 1 warm-up step, "
                      f"{secs:.1f} s"}


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


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--preset", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "fp32-kahan", "fp32-wide"], default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain copy of this many MiB (profiler counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the step shards by x slabs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    res = gpu_run(args)
    scene = res["scene"]
    pse = scene.params.heat == 2
    dom, roof = roofline(res, args.precision, pse)
    for k, v in sorted(res["kernels"].items(), key=lambda kv: -kv[1]["total_ms"]):
        log(f"  {k:20s} {v['launches']:6d} launches  avg {v['avg_ms']:.4f} ms  total {v['total_ms']:.1f} ms")
    log(f"GPU: {res['value']:.4e} particle-steps/s over {args.steps} steps ({res['secs']:.3f} s)")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args)
        log(f"CPU oracle: {cpu['value']:.4e} particle-steps/s on {cpu['cores']} threads")
    out = {
        "metric": "particle-steps/s at 1M particles (FP64 method, single grain, PSE + thermal contact)",
        "value": res["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["secs"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded lattice + jitter, synthetic grain mesh)",
        "config": {"workload": f"orthogonal cut preset {args.preset} (scale {args.scale})",
                   "particles": res["n_particles"], "active_particles": res["n_active"],
                   "precision": args.precision, "seed": args.seed,
                   "pairs_per_particle": res["pairs"] / max(res["n_active"], 1),
                   "candidates_per_particle": res["cand"] / max(res["n_active"], 1)},
        "roofline": roof,
        "kernels": res["kernels"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
