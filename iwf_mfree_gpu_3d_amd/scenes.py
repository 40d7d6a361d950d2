"""Orthogonal-cut workloads built by the library's set-up code (mfg_scene_*), and the
set-up of a device simulation from them, in the manner of the reference's setup_* functions
(src/Benchmarks/benchmarks_single_grain.cu:674-876)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, api
from ._lib import MFG_FP64, check


class Scene:
    """Host arrays and constants of one workload (no device objects)."""

    def __init__(self, preset: int = 2, scale: float = 1.0, seed: int = 1, **overrides):
        L = _lib.load()
        cfg = _lib.cut_config()
        check(L.mfg_scene_preset(int(preset), float(scale), int(seed), C.byref(cfg)))
        for k, v in overrides.items():
            if not hasattr(cfg, k):
                raise AttributeError(f"cut_config has no field {k}")
            setattr(cfg, k, v)
        self.cfg = cfg
        h = L.mfg_scene_build(C.byref(cfg))
        if not h:
            raise _lib.MfgError(_lib.MFG_ERR_ARG, "scene build failed")
        try:
            p = _lib.scene_params()
            check(L.mfg_scene_get_params(h, C.byref(p)))
            self.params = p
            n = p.n_particles
            self.pos = np.zeros((n, 3))
            self.vel = np.zeros((n, 3))
            self.rho, self.T, self.h = np.zeros(n), np.zeros(n), np.zeros(n)
            self.fixed, self.tool_p = np.zeros(n), np.zeros(n)
            dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
            check(L.mfg_scene_get_particles(h, dp(self.pos), dp(self.vel), dp(self.rho), dp(self.T), dp(self.h),
                                            dp(self.fixed), dp(self.tool_p)))
            self.tools = []
            for k in range(p.n_tools):
                nv, nt = C.c_int(0), C.c_int(0)
                check(L.mfg_scene_tool_size(h, k, C.byref(nv), C.byref(nt)))
                verts = np.zeros((nv.value, 3))
                tris = np.zeros((nt.value, 3), dtype=np.int32)
                check(L.mfg_scene_tool_mesh(h, k, dp(verts), tris.ctypes.data_as(C.POINTER(C.c_int))))
                self.tools.append((verts, tris))
        finally:
            L.mfg_scene_free(h)

    @property
    def n(self) -> int:
        return int(self.params.n_particles)

    def bbmin(self):
        return np.array(self.params.bbmin[:])

    def bbmax(self):
        return np.array(self.params.bbmax[:])


def setup_cut(scene: Scene, precision: int = MFG_FP64, single_sqrt: bool = True, max_neighbours: int = 0,
              record_forces: bool = True, pair_pass: int = 0, pair_strip: int = 0):
    """Constants, tools, grid, blanking and particles of `scene` on the device; sets the process
    globals like the reference set-ups do.  Returns (particles, grid)."""
    p = scene.params
    api.reset_globals()
    api.set_options(cspm=bool(p.cspm), heat=int(p.heat), single_sqrt=single_sqrt, max_neighbours=max_neighbours,
                    pair_pass=pair_pass, pair_strip=pair_strip)
    api.actions_setup_corrector_constants(p.corr)
    api.actions_setup_physical_constants(p.phys)
    api.actions_setup_johnson_cook_constants(p.joco)
    api.actions_setup_thermal_constants_wp(p.trml_wp)
    api.interactions_setup_corrector_constants(p.corr)
    api.interactions_setup_physical_constants(p.phys)
    api.interactions_setup_thermal_constants_workpiece(p.trml_wp)
    tools = []
    for verts, tris in scene.tools:
        t = api.tool_3d_gpu(verts, tris, vel=np.array(p.tool_vel[:]), phys=p.phys, alpha=p.contact_alpha)
        t.set_algorithm_type(p.algo)
        t.set_mu(p.mu)
        tools.append(t)
    if p.trml_tool.alpha != 0.0:
        api.actions_setup_thermal_constants_tool(p.trml_tool)
        api.interactions_setup_thermal_constants_tool(p.trml_tool, tools[0] if tools else None)
    api.global_tool = tools
    api.global_time_dt = p.dt
    api.global_time_final = p.dt * p.steps
    particles = api.particle_gpu(scene.pos, scene.vel, scene.rho, scene.T, scene.h, scene.fixed, scene.tool_p,
                                 precision=precision)
    grid = api.grid_gpu_green(scene.n, scene.bbmin(), scene.bbmax(), p.h)
    api.global_blanking = api.blanking(scene.bbmin(), scene.bbmax(), max_vel=p.max_vel)
    if record_forces and tools:
        api.global_tool_forces = api.tool_forces(len(tools))
    return particles, grid


def run(stepper: api.leap_frog, particles, grid, n_steps: int, record_forces: bool = False) -> None:
    """The driver loop of src/mfree_gpu_3d.cpp:155-205 for `n_steps` calls of step(): each step is
    followed by global_time_current += dt."""
    for _ in range(n_steps):
        stepper.step(particles, grid, record_forces)
        api.global_time_step += 1
        api.global_time_current += api.global_time_dt


def steady_tool_temperatures(samples: np.ndarray, steady: np.ndarray, dz: float, hdx: float,
                             T_ref: float) -> np.ndarray:
    """Warm-start temperatures of tool particles from a steady-state solution
    (src/Benchmarks/benchmarks_single_grain.cu:303-373): the steady samples (x, y, z, T) are
    shifted so that their bounding-box minimum meets the tool samples' minimum, interpolated onto
    the tool samples with weights dx_s^3 and smoothing length max(1.5 dx_s, hdx dz), dx_s the
    largest coordinate difference of the first two steady samples, and clamped below at T_ref.
    The reference assigns each tool particle the value of the next sample in its list (the
    increment of its iterator comes first); here every particle gets its own value."""
    steady = np.asarray(steady, dtype=np.float64).reshape(-1, 4)
    samples = np.asarray(samples, dtype=np.float64).reshape(-1, 3)
    dx_s = float(np.max(np.abs(steady[0, :3] - steady[1, :3])))
    shifted = steady.copy()
    shifted[:, :3] = steady[:, :3] - steady[:, :3].min(axis=0) + samples.min(axis=0)
    T = api.interp_temps(shifted, dx_s ** 3, max(1.5 * dx_s, hdx * dz), samples)
    return np.maximum(T, T_ref)
