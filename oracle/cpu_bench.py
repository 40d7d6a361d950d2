"""CPU baseline of the step: the oracle (a CPU restatement of the reference's step, test
infrastructure) timed on the host cores on a bounded sample of a workload.  Run as
``python -m oracle.cpu_bench --preset 2 --steps 8 --threads 16``; prints one JSON object.

The caller sets OMP_NUM_THREADS and OMP_PROC_BIND=close in this process's environment, so the
OpenMP runtime starts with them; MFG_ORACLE_LIB names the library to time (the native build).
Timed with a steady clock around the step loop only (set-up, upload and one warm-up step
excluded); particle-steps/s counts the active particles after blanking.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--preset", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--threads", type=int, default=1)
    a = ap.parse_args()
    import helpers
    from iwf_mfree_gpu_3d_amd import scenes
    scene = scenes.Scene(a.preset, a.scale, a.seed)
    sim = helpers.oracle_from_scene(scene, single_sqrt=True, threads=a.threads)
    dt = scene.params.dt
    sim.set_time(dt, dt)
    sim.run(1)  # warm-up: the first sort starts from the unsorted lattice
    t0 = time.perf_counter()
    sim.run(a.steps)
    secs = time.perf_counter() - t0
    n = sim.counts()[0]
    print(json.dumps({"particles": scene.n, "active": n, "steps": a.steps, "seconds": secs,
                      "threads": a.threads, "value": n * a.steps / secs}))


if __name__ == "__main__":
    main()
