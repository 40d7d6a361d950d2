"""Autotuner of the step's launch choices, per size bucket (batched and served forms).

Every choice offered leaves results bit-identical, so the tuner only measures time:
- the pair pass's form (MFG_PAIR_LIST, one thread per particle, or MFG_PAIR_TILED, one
  wavefront per grid column) and the list pass's visiting order (strip width in grid columns, or
  the sorted order).  The tiled form is offered in FP64 only: both forms form the same sums in
  the same order, but in the FP32 modes the two kernels' FMA contraction can differ at rounding
  level;
- the neighbour pass's cell tile (consecutive grid cells per wavefront: 1, 2, 4, 8, 16), which
  decides how much an empty cell of a grid with wide margins costs;
- the list pair pass's workgroup size (64, 128, 256 threads): every wavefront works alone, and
  smaller workgroups spread a small instance over more compute units.

For instances on a stream pool (serve mode "streams"), `tune` picks per (bucket, precision) the
form and order first, then the cell tile with them, then the pair pass's workgroup size.  For batches (serve mode "batched") the cell
tile is a property of the whole launch: `tune_batched` measures a throwaway batch of the first
requests of each precision with every tile.

The re-sort interval is offered only when the caller opts in (`sort_intervals`): the reference
rebuilds its cell lists from a fresh sort every step (src/leap_frog.cu:139-145), and sorting on
every k-th step only changes the neighbour lists and so the results (the lists are bit-exact with
the reference's only at interval 1).  Opted in, the last stage measures each interval with the
chosen form, order and tile, and keeps one only if it is faster.  Workgroup sizes and waves per
SIMD of the other kernels are compile-time constants (the measured alternatives are documented
with them).

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
# cells per wavefront of the neighbour pass; 0 the library's automatic choice
TILES = (1, 2, 4, 8, 16)
# threads per workgroup of the list pair pass; 0 the library's default (256)
BLOCKS = (64, 128, 256)
# (pair pass, strip, cell tile, sort interval, pair workgroup); sort interval 0: every step
DEFAULT = (api.MFG_PAIR_LIST, 0, 0, 0, 0)


def candidates_for(precision: int, candidates=CANDIDATES) -> tuple:
    """The candidates whose results are bit-identical to the default's in `precision`."""
    return tuple(c for c in candidates if c[0] == api.MFG_PAIR_LIST or precision == api.MFG_FP64)


def bucket(n: int) -> int:
    return int(math.floor(math.log2(max(int(n), 1))))


class Table:
    """(bucket, precision) -> (pair pass, strip, cell tile, sort interval, pair workgroup), with
    the measured seconds per step; precision -> cell tile of a batch.  Shorter choices (tables
    written before the later entries) read with 0 (the default) for what they lack."""

    def __init__(self):
        self.choice = {}
        self.times = {}
        self.batch_tile = {}
        self.batch_times = {}

    def get(self, n: int, precision: int) -> tuple:
        b = bucket(n)
        tuned = [k for k in self.choice if k[1] == precision]
        if not tuned:
            return DEFAULT
        key = min(tuned, key=lambda k: (abs(k[0] - b), k[0]))
        return self.choice[key]

    def to_json(self) -> str:
        return json.dumps({"choice": [[b, p, *c] for (b, p), c in sorted(self.choice.items())],
                           "times": [[b, p, *c, t] for (b, p, c), t in sorted(self.times.items())],
                           "batch_tile": [[p, t] for p, t in sorted(self.batch_tile.items())],
                           "batch_times": [[p, t, s] for (p, t), s in sorted(self.batch_times.items())]})

    @staticmethod
    def from_json(text: str) -> "Table":
        d = json.loads(text)
        t = Table()
        for row in d["choice"]:
            t.choice[(row[0], row[1])] = (tuple(row[2:]) + (0, 0))[:5]
        for row in d.get("times", []):
            t.times[(row[0], row[1], (tuple(row[2:-1]) + (0, 0))[:5])] = row[-1]
        for p, tile in d.get("batch_tile", []):
            t.batch_tile[p] = tile
        for p, tile, s in d.get("batch_times", []):
            t.batch_times[(p, tile)] = s
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


def tune(requests, candidates=CANDIDATES, tiles=TILES, steps: int = 6, warmup: int = 2, measure=measure,
         log=None, sort_intervals=(), blocks=BLOCKS) -> Table:
    """One representative request per (bucket, precision) among `requests` (the first seen):
    every form / order candidate measured on it with the automatic cell tile, then every cell
    tile with the fastest of those, then every pair-pass workgroup size, then (opt-in, results
    change) every re-sort interval in `sort_intervals` with the fastest of all the others; ties
    go to the earlier candidate."""
    table = Table()
    reps = {}
    for r in requests:
        nx, ny, nz = r.lattice()
        reps.setdefault((bucket(nx * ny * nz), r.precision), r)
    for key, r in sorted(reps.items()):
        best = None
        for cand in candidates_for(key[1], candidates):
            full = (cand[0], cand[1], 0, 0, 0)
            s = measure(r, full, steps=steps, warmup=warmup)
            table.times[(key[0], key[1], full)] = s
            if best is None or s < best[0]:
                best = (s, full)
        for tile in tiles:
            full = (best[1][0], best[1][1], tile, 0, 0)
            s = measure(r, full, steps=steps, warmup=warmup)
            table.times[(key[0], key[1], full)] = s
            if s < best[0]:
                best = (s, full)
        for bs in blocks:
            full = (best[1][0], best[1][1], best[1][2], 0, int(bs))
            s = measure(r, full, steps=steps, warmup=warmup)
            table.times[(key[0], key[1], full)] = s
            if s < best[0]:
                best = (s, full)
        for k in sort_intervals:
            if k <= 1:
                continue
            full = (best[1][0], best[1][1], best[1][2], int(k), best[1][4])
            s = measure(r, full, steps=steps, warmup=warmup)
            table.times[(key[0], key[1], full)] = s
            if s < best[0]:
                best = (s, full)
        table.choice[key] = best[1]
        if log:
            log(f"autotune bucket 2^{key[0]} precision {key[1]}: pair pass {best[1][0]}, strip {best[1][1]}, "
                f"cell tile {best[1][2]}, workgroup {best[1][4] or 256}, sort interval {best[1][3] or 1} "
                f"({best[0] * 1e3:.3f} ms/step)")
    return table


def measure_batch(requests, tile: int, steps: int = 4, warmup: int = 2) -> float:
    """Seconds per batched step of fresh instances of `requests` with cell tile `tile`."""
    from .serve import Instance
    b = api.batch(tile)
    insts = []
    try:
        for r in requests:
            inst = Instance(r, 0)
            b.add(inst.particles, inst.grid, inst.tools, inst.blanking, None, dt=inst.dt, t_current=inst.t)
            insts.append(inst)
        b.step(warmup)
        api.synchronize()
        t0 = time.perf_counter()
        b.step(steps)
        api.synchronize()
        return (time.perf_counter() - t0) / steps
    finally:
        del b
        for inst in insts:
            inst.release()
        del insts
        gc.collect()


def tune_batched(requests, tiles=TILES, sample: int = 16, steps: int = 4, warmup: int = 2,
                 measure=measure_batch, log=None, table: Table | None = None) -> Table:
    """The cell tile of each precision's batch: the first `sample` requests of the precision as a
    throwaway batch, measured with every tile (ties: the smaller tile)."""
    table = table or Table()
    by_prec = {}
    for r in requests:
        by_prec.setdefault(r.precision, []).append(r)
    for prec, reqs in sorted(by_prec.items()):
        best = None
        for tile in tiles:
            s = measure(reqs[:sample], tile, steps=steps, warmup=warmup)
            table.batch_times[(prec, tile)] = s
            if best is None or s < best[0]:
                best = (s, tile)
        table.batch_tile[prec] = best[1]
        if log:
            log(f"autotune batch precision {prec}: cell tile {best[1]} ({best[0] * 1e3:.3f} ms per batched step of "
                f"{min(sample, len(reqs))} instances)")
    return table
