"""Shared helpers of the test suite: the same workload on the CPU oracle and on the device."""
from __future__ import annotations

import numpy as np

import oracle as O
from iwf_mfree_gpu_3d_amd import scenes


def oracle_from_scene(scene: scenes.Scene, single_sqrt: bool = True, threads: int = 0) -> O.OracleSim:
    """Configure the CPU oracle exactly as scenes.setup_cut configures the device."""
    p = scene.params
    sim = O.OracleSim(scene.pos, scene.vel, scene.rho, scene.T, scene.h, scene.fixed, scene.tool_p)
    if threads:
        sim.set_threads(threads)
    sim.set_options(cspm=bool(p.cspm), heat=int(p.heat), single_sqrt=single_sqrt)
    phys = O.Phys(*[getattr(p.phys, f) for f, _ in O.Phys._fields_])
    corr = O.Corr(*[getattr(p.corr, f) for f, _ in O.Corr._fields_])
    joco = O.Joco(*[getattr(p.joco, f) for f, _ in O.Joco._fields_])
    twp = O.Trml(*[getattr(p.trml_wp, f) for f, _ in O.Trml._fields_])
    ttool = O.Trml(*[getattr(p.trml_tool, f) for f, _ in O.Trml._fields_]) if p.trml_tool.alpha != 0.0 else None
    sim.setup_constants(phys, corr, joco, twp, ttool)
    sim.set_grid_locked(scene.bbmin(), scene.bbmax(), p.h)
    sim.set_grid_params()
    sim.set_blanking(scene.bbmin(), scene.bbmax(), max_vel_squared=p.max_vel * p.max_vel)
    for verts, tris in scene.tools:
        sim.add_tool(verts, tris, vel=np.array(p.tool_vel[:]), mass=p.phys.mass, alpha=p.contact_alpha,
                     algo=int(p.algo), mu=p.mu)
    return sim
