"""Batched and served forms of the cutting step: many independent cutting instances on one GPU.

A request is one orthogonal cut (block of 4:2:1 lattice cells, wedge tool) with its own particle
count, depth of cut, speed, rake angle, Johnson-Cook constants, precision mode and number of steps.
The request loop admits a request once it has arrived and the HBM budget has room for its state,
keeps the state device-resident until its last step, and shares the GPU between the admitted
instances by continuous batching: every round advances each admitted instance by a few steps,
and a request that arrives meanwhile joins at the next round.  Each request's time to result is
its completion time minus its arrival time.

Instances share no device state.  Two ways to run them:
- "batched" (default): the admitted instances of one precision form a batch (mfg_batch_*), so every
  kernel of the step is launched once over all of them, each with its own grid, constants, tools
  and clock; a round advances a batch by up to `steps_per_round` steps (no further than its
  nearest finishing instance), optionally replayed from a captured graph.  Each precision's batch
  has its own stream.  Instances join and leave between rounds.
- "streams": each instance queues its own steps on one stream of a small pool (the least loaded at
  admission).  The library's constants are process-wide (as the reference's are), so an instance
  re-installs its own constants, tools, blanking box and clock before each of its steps.
Either way an instance computes exactly what it computes alone, so a result does not depend on the
mix of requests it ran with.  The batched configuration is the same loop with every request
arriving at time zero.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import time

import numpy as np

from . import _lib, api, scenes


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float          # s after the start of the run
    n_target: int           # particles asked for (the lattice rounds it)
    depth: float            # cm
    speed: float            # cm/us
    rake_deg: float
    jc_scale: tuple         # factors on Johnson-Cook A, B, C
    precision: int          # MFG_FP64 / MFG_FP32_KAHAN / MFG_FP32_WIDE
    steps: int
    seed: int

    def lattice(self):
        """4:2:1 block: nz = ceil((N / 8)^(1/3)), ny = 2 nz, nx = 4 nz."""
        nz = max(2, int(math.ceil((self.n_target / 8.0) ** (1.0 / 3.0))))
        return 4 * nz, 2 * nz, nz


def sweep_requests(count: int | None = None, rate: float = 4.0, duration: float = 600.0, seed: int = 1,
                   n_min: float = 2e4, n_max: float = 2e6, steps=(100, 1000), fp32_share: float = 0.7,
                   batched: bool = False) -> list:
    """Requests of a parameter sweep.  Served form: Poisson arrivals at `rate` per second over
    `duration` seconds (or `count` requests); batched form (`batched`): `count` requests at time
    zero.  Particle counts log-uniform in [n_min, n_max]; depth U[5e-4, 4e-3] cm; speed
    U[5e-4, 5e-3] cm/us; rake U[-30, 10] degrees; A, B, C each scaled by U(0.9, 1.1);
    `fp32_share` of the requests in FP32 with Kahan sums, the rest FP64; steps uniform in the
    closed range `steps`."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    k = 0
    while True:
        if batched:
            if k >= (count or 0):
                break
            arrival = 0.0
        else:
            t += rng.exponential(1.0 / rate)
            if (count is not None and k >= count) or (count is None and t > duration):
                break
            arrival = t
        n = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        prec = _lib.MFG_FP32_KAHAN if rng.uniform() < fp32_share else _lib.MFG_FP64
        out.append(Request(rid=k, arrival=arrival, n_target=int(n), depth=float(rng.uniform(5e-4, 4e-3)),
                           speed=float(rng.uniform(5e-4, 5e-3)), rake_deg=float(rng.uniform(-30.0, 10.0)),
                           jc_scale=tuple(float(x) for x in rng.uniform(0.9, 1.1, 3)), precision=prec,
                           steps=int(rng.integers(steps[0], steps[1] + 1)), seed=int(seed * 100003 + k)))
        k += 1
    return out


# device bytes per particle of one instance, from the library's allocations: double-buffered
# state 656, pair record 192 (FP64) / 112 (FP32), nl4 + pse2 48, CSPM entries 32, neighbour sums
# 200, 256-entry neighbour list 1024, keys / order / counts / visiting order 20; plus cell tables
# and tools
STATE_BYTES = {_lib.MFG_FP64: 2172.0, _lib.MFG_FP32_KAHAN: 2092.0, _lib.MFG_FP32_WIDE: 2092.0}


def device_bytes(req: Request) -> float:
    nx, ny, nz = req.lattice()
    n = nx * ny * nz
    return STATE_BYTES[req.precision] * n * 1.05 + 32e6


class Instance:
    """The device state of one request and the constants it runs with."""

    DX = 5e-4  # lattice spacing [cm]

    def __init__(self, req: Request, stream: int = 0, choice: tuple = (api.MFG_PAIR_LIST, 0, 0, 0, 0),
                 max_neighbours: int = 0):
        self.req = req
        self.stream = stream
        # (pair pass, strip, cell tile, sort interval, pair workgroup) the instance's steps run with
        self.choice = (tuple(choice) + (0, 0, 0, 0, 0))[:5]
        api.select_stream(stream)
        nx, ny, nz = req.lattice()
        cfg = _lib.cut_config()
        _lib.check(_lib.load().mfg_scene_preset(1, 1.0, req.seed, C.byref(cfg)))
        self.scene = scenes.Scene(1, 1.0, req.seed, nx=nx, ny=ny, nz=nz, dx=self.DX, depth_of_cut=req.depth,
                                  speed=req.speed, rake_deg=req.rake_deg, jc_A=cfg.jc_A * req.jc_scale[0],
                                  jc_B=cfg.jc_B * req.jc_scale[1], jc_C=cfg.jc_C * req.jc_scale[2],
                                  steps=req.steps)
        self.particles, self.grid = scenes.setup_cut(self.scene, precision=req.precision, record_forces=False,
                                                     max_neighbours=max_neighbours)
        self.tools = api.global_tool
        self.blanking = api.global_blanking
        self.dt = api.global_time_dt
        self.t = self.dt
        self.k = 0
        self.stepper = api.leap_frog(self.scene.n, self.grid.num_cell())
        self.n = self.scene.n

    def install(self) -> None:
        """The library's constants and the driver globals of this instance."""
        p = self.scene.params
        api.constants_reset()
        api.set_options(cspm=bool(p.cspm), heat=int(p.heat), single_sqrt=True,
                        pair_pass=int(self.choice[0]), pair_strip=int(self.choice[1]),
                        cell_tile=int(self.choice[2]), sort_interval=int(self.choice[3]),
                        pair_block=int(self.choice[4]))
        api.actions_setup_corrector_constants(p.corr)
        api.actions_setup_physical_constants(p.phys)
        api.actions_setup_johnson_cook_constants(p.joco)
        api.actions_setup_thermal_constants_wp(p.trml_wp)
        api.interactions_setup_corrector_constants(p.corr)
        api.interactions_setup_physical_constants(p.phys)
        api.interactions_setup_thermal_constants_workpiece(p.trml_wp)
        if p.trml_tool.alpha != 0.0:
            api.actions_setup_thermal_constants_tool(p.trml_tool)
            api.interactions_setup_thermal_constants_tool(p.trml_tool, self.tools[0] if self.tools else None)
        api.global_tool = self.tools
        api.global_blanking = self.blanking
        api.global_tool_forces = None
        api.global_time_dt = self.dt
        api.global_time_current = self.t

    def advance(self, steps: int) -> int:
        """Up to `steps` steps (fewer at the end of the request); returns the number taken."""
        k = min(steps, self.req.steps - self.k)
        if k <= 0:
            return 0
        api.select_stream(self.stream)
        self.install()
        for _ in range(k):
            self.stepper.step(self.particles, self.grid)
            self.t += self.dt
            api.global_time_current = self.t
        self.k += k
        return k

    @property
    def done(self) -> bool:
        return self.k >= self.req.steps

    def finish(self) -> int:
        """Waits for the instance's queued work (its stream selected for what follows) and reads
        its device error counters, which keep the first failure of any step: MFG_OK, or the code
        of a neighbour-list overflow (MFG_ERR_CAPACITY), an unblanked particle outside the grid
        (MFG_ERR_GRID) or every particle blanked (MFG_ERR_ALL_BLANKED)."""
        api.select_stream(self.stream)
        api.synchronize()
        try:
            self.particles.counts()
        except _lib.MfgError as e:
            return e.code
        return _lib.MFG_OK

    def release(self) -> None:
        api.select_stream(self.stream)
        self.particles = self.grid = self.tools = self.blanking = self.stepper = None


class _Group:
    """The batch of one precision mode and its stream."""

    def __init__(self, stream: int, cell_tile: int = 0):
        self.stream = stream
        api.select_stream(stream)
        self.batch = api.batch(cell_tile)
        self.members = []  # Instance per batch slot


def serve(requests: list, hbm_budget: float = 250e9, steps_per_round: int = 5, on_done=None,
          clock=time.perf_counter, log=None, streams: int = 4, tuning=None, max_neighbours: int = 0,
          mode: str = "batched", use_graph: bool = True, cell_tile: int = 0) -> dict:
    """Runs the request loop until every request has finished.  `on_done(instance)` is called
    for each instance that finished without an error, before its state is released; a request
    whose instance reports a device error gets that code as its record's "error" and no time to
    result (the error counters are sticky, so the check at the end covers every step).
    `max_neighbours`: neighbour-list capacity of the instances (0: the library default).
    `mode`: "batched" or "streams" (see the module notes); `use_graph`: batched rounds replay a
    captured step; `cell_tile`: cells per wavefront of the batches' neighbour pass (0 automatic).
    `streams`: size of the stream pool the instances are spread over.  `tuning`: an
    autotune.Table; "streams": each instance runs with its bucket's (pair pass, strip, cell tile),
    "batched": each precision's batch with its cell tile (None: the library defaults).  Returns
    per-request records and p50 / p99 time to result and the aggregate particle-steps/s."""
    if mode not in ("batched", "streams"):
        raise ValueError(f"mode {mode!r}")
    _lib.init()
    api.stream_pool(streams)
    load = [0] * streams  # active instances per stream
    groups = {}  # precision -> _Group (batched mode)
    pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
    active = []
    used = 0.0
    records = {}
    t0 = clock()
    steps_total = 0
    setup_s = 0.0  # time spent setting up admitted instances (inside the makespan)
    while pending or active:
        now = clock() - t0
        while pending and pending[0].arrival <= now:
            need = device_bytes(pending[0])
            if active and used + need > hbm_budget:
                break  # first come, first served: later requests wait behind this one
            r = pending.pop(0)
            ts = clock()
            if mode == "batched":
                grp = groups.get(r.precision)
                if grp is None:
                    tile = tuning.batch_tile.get(r.precision, cell_tile) if tuning is not None else cell_tile
                    grp = groups[r.precision] = _Group(len(groups) % streams, tile)
                inst = Instance(r, grp.stream, max_neighbours=max_neighbours)
                grp.batch.add(inst.particles, inst.grid, inst.tools, inst.blanking, None, dt=inst.dt,
                              t_current=inst.t)
                grp.members.append(inst)
            else:
                k = min(range(streams), key=lambda q: load[q])
                load[k] += 1
                nx, ny, nz = r.lattice()
                choice = (tuning.get(nx * ny * nz, r.precision) if tuning is not None
                          else (api.MFG_PAIR_LIST, 0, cell_tile))
                inst = Instance(r, k, choice, max_neighbours=max_neighbours)
            used += need
            setup_s += clock() - ts
            active.append(inst)
            records[r.rid] = {"rid": r.rid, "arrival": r.arrival, "admitted": clock() - t0, "particles": inst.n,
                              "steps": r.steps, "precision": r.precision}
            if log:
                log(f"t={clock() - t0:8.2f}s admit request {r.rid}: {inst.n} particles, {r.steps} steps, "
                    f"{len(active)} active")
        if not active:
            if pending:
                time.sleep(max(0.0, min(0.05, pending[0].arrival - (clock() - t0))))
            continue
        if mode == "batched":
            for grp in groups.values():
                if not grp.members:
                    continue
                api.select_stream(grp.stream)
                k = min(steps_per_round, min(i.req.steps - i.k for i in grp.members))
                grp.batch.step(k, use_graph=use_graph)
                for i in grp.members:
                    i.k += k
                    steps_total += i.n * k
        else:
            for inst in active:
                steps_total += inst.n * inst.advance(steps_per_round)
        for inst in list(active):
            if inst.done:
                err = inst.finish()
                rec = records[inst.req.rid]
                rec["finished"] = clock() - t0
                if err in (None, _lib.MFG_OK):
                    rec["time_to_result"] = rec["finished"] - rec["arrival"]
                    if on_done:
                        on_done(inst)
                else:
                    rec["error"] = int(err)
                    if log:
                        log(f"t={rec['finished']:8.2f}s request {inst.req.rid} failed with error {err}")
                if mode == "batched":
                    grp = groups[inst.req.precision]
                    slot = grp.members.index(inst)
                    api.select_stream(grp.stream)
                    grp.batch.remove(slot)
                    grp.members.pop(slot)
                else:
                    load[inst.stream] -= 1
                inst.release()
                active.remove(inst)
                used -= device_bytes(inst.req)
    api.device_synchronize()
    api.select_stream(0)
    makespan = clock() - t0
    ttr = np.array([r["time_to_result"] for r in records.values() if "time_to_result" in r])
    return {"requests": [records[k] for k in sorted(records)],
            "p50_s": float(np.percentile(ttr, 50)) if ttr.size else 0.0,
            "p99_s": float(np.percentile(ttr, 99)) if ttr.size else 0.0,
            "particle_steps": steps_total, "makespan_s": makespan, "setup_s": setup_s,
            "particle_steps_per_s": steps_total / makespan if makespan > 0 else 0.0}
