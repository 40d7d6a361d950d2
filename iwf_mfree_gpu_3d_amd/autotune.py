"""Autotuner of the pair pass's run-time choices, per size bucket (batched and served forms).

The choices are the pair-pass form (MFG_PAIR_LIST, one thread per particle, or MFG_PAIR_TILED,
one wavefront per grid column) and the list pass's visiting order (strip width in grid columns,
or the sorted order).  Only choices that leave results bit-identical are offered: the visiting
order in every precision, the tiled form in FP64 only (both forms form the same sums in the same
order, but in the FP32 modes the two kernels' FMA contraction can differ at rounding level).  So
the tuner only measures time.

Not a choice here: the re-sort interval.  The reference rebuilds its cell lists from a fresh
sort every step (src/leap_frog.cu:139-145), and sorting less often would change which
particles are neighbours.  Block and cell-tile sizes are compile-time constants of the kernels.

Buckets are powers of two of the particle count; an instance takes the choice of its own
bucket, else of the nearest tuned one, else the library default.
"""
from __future__ import annotations

import gc
import json
import math
import time

from . import api

# (pair pass, strip): 0 the library's default, -1 the sorted order
CANDIDATES = ((api.MFG_PAIR_LIST, 0), (api.MFG_PAIR_LIST, -1), (api.MFG_PAIR_LIST, 8),
              (api.MFG_PAIR_LIST, 16), (api.MFG_PAIR_TILED, 0))
DEFAULT = (api.MFG_PAIR_LIST, 0)


def candidates_for(precision: int, candidates=CANDIDATES) -> tuple:
    """The candidates whose results are bit-identical to the default's in `precision`."""
    return tuple(c for c in candidates if c[0] == api.MFG_PAIR_LIST or precision == api.MFG_FP64)


def bucket(n: int) -> int:
    return int(math.floor(math.log2(max(int(n), 1))))


class Table:
    """(bucket, precision) -> (pair pass, strip), with the measured seconds per step."""

    def __init__(self):
        self.choice = {}
        self.times = {}

    def get(self, n: int, precision: int) -> tuple:
        b = bucket(n)
        tuned = [k for k in self.choice if k[1] == precision]
        if not tuned:
            return DEFAULT
        key = min(tuned, key=lambda k: (abs(k[0] - b), k[0]))
        return self.choice[key]

    def to_json(self) -> str:
        return json.dumps({"choice": [[b, p, c[0], c[1]] for (b, p), c in sorted(self.choice.items())],
                           "times": [[b, p, c[0], c[1], t] for (b, p, c), t in sorted(self.times.items())]})

    @staticmethod
    def from_json(text: str) -> "Table":
        d = json.loads(text)
        t = Table()
        for b, p, pp, st in d["choice"]:
            t.choice[(b, p)] = (pp, st)
        for b, p, pp, st, s in d.get("times", []):
            t.times[(b, p, (pp, st))] = s
        return t

    def save(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(self.to_json())

    @staticmethod
    def load(path: str) -> "Table":
        with open(path) as f:
            return Table.from_json(f.read())


def measure(req, candidate: tuple, steps: int = 6, warmup: int = 2) -> float:
    """Seconds per step of a fresh instance of `req` run with `candidate` (device-synchronised).
    The instance's device memory is freed before returning: left to the cyclic garbage collector,
    its frees would come during the request loop that follows, and every free waits for the whole
    device (the batched form ran 23 % slower so)."""
    from .serve import Instance
    inst = Instance(req, 0, choice=candidate)
    try:
        inst.advance(warmup)
        inst.finish()
        t0 = time.perf_counter()
        k = inst.advance(steps)
        inst.finish()
        return (time.perf_counter() - t0) / max(k, 1)
    finally:
        inst.release()
        del inst
        gc.collect()


def tune(requests, candidates=CANDIDATES, steps: int = 6, warmup: int = 2, measure=measure,
         log=None) -> Table:
    """One representative request per (bucket, precision) among `requests` (the first seen),
    each candidate measured on it; the fastest is the bucket's choice (ties: the earlier
    candidate)."""
    table = Table()
    reps = {}
    for r in requests:
        nx, ny, nz = r.lattice()
        reps.setdefault((bucket(nx * ny * nz), r.precision), r)
    for key, r in sorted(reps.items()):
        best = None
        for cand in candidates_for(key[1], candidates):
            s = measure(r, cand, steps=steps, warmup=warmup)
            table.times[(key[0], key[1], tuple(cand))] = s
            if best is None or s < best[0]:
                best = (s, tuple(cand))
        table.choice[key] = best[1]
        if log:
            log(f"autotune bucket 2^{key[0]} precision {key[1]}: pair pass {best[1][0]}, strip {best[1][1]} "
                f"({best[0] * 1e3:.3f} ms/step)")
    return table
