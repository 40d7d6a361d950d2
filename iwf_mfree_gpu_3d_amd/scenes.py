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


def single_grain_1tool(verts: np.ndarray, tris: np.ndarray, feed_particles: float = 18.0,
                       realscale: bool = False, advance: float = 0.0) -> Scene:
    """The reference's single-grain runs around a shipped tool mesh: its default run,
    setup_single_grain_1tool (src/Benchmarks/benchmarks_single_grain.cu:674-876), with the pin mesh
    pin25_reduziert_20062018.k in cm; or, with `realscale`, setup_single_grain_1tool_realscale
    (:473-672) with the diamond Diamant_von_Mansur_04072018_SDB1125-2025-D851.k in microns.

    Default run: a 160 x 160 um block, three feeds (60 um) deep, `feed_particles` particle
    layers across the 20 um feed (18 in the reference: 1,073,574 particles), h = 1.7 dz; the tool
    lifted to z = 0, pushed down by one feed, centred in y and moved to reach 6.4 um past the
    block's x = 0 face; Johnson-Cook A 862 MPa, B 331 MPa, C 0.01, m 0.8, n 0.34, T_melt 1836 K;
    the bottom layer and the last x layer fixed; grid and blanking box up to 3 x the tool's top.
    Realscale: the mesh scaled from microns to cm, a 30 um feed, 540 x 640 um block (5 layers per
    feed in the reference: 123,255 particles), the tool reaching 260 um past x = 0, Johnson-Cook
    A 1104 MPa, B 1036 MPa, C 0.0139, m 0.7794, n 0.6349, T_melt 1723 K (Akbari 2016), the last
    five x layers fixed, boxes up to 2 x the tool's top.
    Both: Brookshaw heat, no CSPM (the reference's compile-time defaults, src/types.h:102-105),
    tool speed 30 m/s along x, friction 0.35, exhaustive contact, blanking at speeds above
    4 cm/us (the constructor squares it).
    `advance`: distance [cm] the tool has already travelled along x when the run starts (0: the
    reference's set-up; engagement_advance gives the distance at which the tool's lower front
    reaches the block, thousands of steps into the reference's run).  Returns a Scene (no device
    objects)."""
    verts = np.asarray(verts, dtype=np.float64).copy()
    tris = np.asarray(tris, dtype=np.int32)
    if realscale:
        verts = verts * (1.0 / 10000.0)   # micron to cm
    target_feed = (30 if realscale else 20) * 1e-6 * 100
    depth_fac = 3
    depth_wp = depth_fac * target_feed
    nz = int(depth_fac * feed_particles)
    dz = depth_wp / (nz - 1)
    bmin, bmax = verts.min(0), verts.max(0)
    verts[:, 2] += -bmin[2]
    verts[:, 2] += -target_feed
    ly_tool = bmax[1] - bmin[1]
    verts[:, 1] += -bmin[1] - ly_tool / 2.0
    nudge = 260 * 100 * 1e-6 if realscale else 6.4e-4
    lx_tool = bmax[0] - bmin[0]
    verts[:, 0] += -bmin[0] - lx_tool + nudge
    bmax_tool = verts.max(0)
    verts[:, 0] += advance
    if realscale:
        lx, ly = 2 * 270 * 1e-6 * 100, 2 * 320 * 1e-6 * 100
    else:
        lx = ly = 2 * 80 * 1e-6 * 100
    lz = depth_wp
    nx, ny = int(lx / dz), int(ly / dz)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    pos = np.stack([g[:, 0] * dz, g[:, 1] * dz - ly / 2.0, -g[:, 2] * dz], axis=1)
    n = pos.shape[0]
    hdx = 1.7
    p = _lib.scene_params()
    E, nu, rho0 = 1.1, 0.35, 4.43
    G = E / (2.0 * (1.0 + nu))
    p.phys.E, p.phys.nu, p.phys.rho0, p.phys.G = E, nu, rho0, G
    p.phys.K = 2.0 * G * (1 + nu) / (3 * (1 - 2 * nu))
    p.phys.mass = dz * dz * dz * rho0
    j = p.joco
    if realscale:
        j.A, j.B, j.C, j.m, j.n = 0.0110400, 0.0103600, 0.0139000, 0.7794, 0.6349
        j.Tref, j.Tmelt, j.eps_dot_ref, j.clamp_temp = 300.0, 1723.0, 1e-6, 1.0
    else:
        j.A, j.B, j.C, j.m, j.n = 0.0086200, 0.0033100, 0.0100000, 0.8, 0.34
        j.Tref, j.Tmelt, j.eps_dot_ref, j.clamp_temp = 300.0, 1836.0, 1e-6, 1.0
    t = p.trml_wp
    t.cp, t.tq, t.k = 553e-8, 0.9, 7.1e-13
    t.alpha = t.k / (rho0 * t.cp)
    t.T_init = j.Tref
    c0 = np.sqrt(p.phys.K / rho0)
    c = p.corr
    c.alpha, c.beta, c.eta, c.xspheps, c.stresseps = 1.0, 1.0, 0.1, 0.5, 0.3
    h1 = 1.0 / (hdx * dz)
    q = dz * h1
    c.wdeltap = (1.0 / np.pi) * h1 * h1 * h1 * (1 - 1.5 * q * q * (1 - 0.5 * q))
    v_tool = 3e-3
    cfl = 0.3
    p.dt = 0.5 * cfl * hdx * dz / (np.sqrt(v_tool) + c0)
    p.steps = int(round(lx / v_tool / p.dt))
    p.tool_vel[0], p.tool_vel[1], p.tool_vel[2] = v_tool, 0.0, 0.0
    p.h = hdx * dz
    bb0 = (-15.0 * dz, -15.0 * dz - ly / 2, -1.5 * nz * dz)
    bb1 = (lx + 15.0 * dz, ly + 15.0 * dz - ly / 2, (2.0 if realscale else 3.0) * bmax_tool[2])
    for k in range(3):
        p.bbmin[k], p.bbmax[k] = bb0[k], bb1[k]
    p.max_vel = 2.0 * 2.0
    p.cspm, p.heat, p.algo = 0, _lib.MFG_HEAT_BROOKSHAW, _lib.MFG_CONTACT_EXHAUSTIVE
    p.mu, p.contact_alpha = 0.35, 0.1
    p.n_tools, p.n_particles, p.n_workpiece = 1, n, n
    sc = Scene.__new__(Scene)
    sc.cfg = None
    sc.params = p
    sc.pos = pos
    sc.vel = np.zeros((n, 3))
    sc.rho = np.full(n, rho0)
    sc.T = np.full(n, j.Tref)
    sc.h = np.full(n, hdx * dz)
    x_fixed = lx - 5 * dz if realscale else lx - dz / 2.0
    sc.fixed = ((pos[:, 2] < -lz + dz / 2.0) | (pos[:, 0] > x_fixed)).astype(np.float64)
    sc.tool_p = np.zeros(n)
    sc.tools = [(verts, tris)]
    return sc


def engagement_advance(scene: Scene, samples: int = 64) -> float:
    """Distance along +x from the tool's front within the block's depth (z in [z_min, z_max] of
    the particles) to the block's x = min face: the tool surface sampled on a barycentric grid
    of every triangle."""
    verts, tris = scene.tools[0]
    z0, z1 = scene.pos[:, 2].min(), scene.pos[:, 2].max()
    u, v = np.meshgrid(np.linspace(0, 1, samples), np.linspace(0, 1, samples))
    keep = u + v <= 1.0
    u, v = u[keep], v[keep]
    P = verts[tris]
    pts = (P[:, None, 0] * (1 - u - v)[None, :, None] + P[:, None, 1] * u[None, :, None]
           + P[:, None, 2] * v[None, :, None]).reshape(-1, 3)
    band = (pts[:, 2] >= z0) & (pts[:, 2] <= z1)
    return float(scene.pos[:, 0].min() - pts[band, 0].max())


def setup_cut(scene: Scene, precision: int = MFG_FP64, single_sqrt: bool = True, max_neighbours: int = 0,
              record_forces: bool = True, pair_pass: int = 0, pair_strip: int = 0, cell_tile: int = 0,
              sort_interval: int = 0, pair_block: int = 0):
    """Constants, tools, grid, blanking and particles of `scene` on the device; sets the process
    globals like the reference set-ups do.  Returns (particles, grid)."""
    p = scene.params
    api.reset_globals()
    api.set_options(cspm=bool(p.cspm), heat=int(p.heat), single_sqrt=single_sqrt, max_neighbours=max_neighbours,
                    pair_pass=pair_pass, pair_strip=pair_strip, cell_tile=cell_tile, sort_interval=sort_interval,
                    pair_block=pair_block)
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
