"""Host-side mirror of the reference call surface, over the C ABI of libmfree_hip.so.

Names, argument meaning and process globals follow the reference so that a driver written
against it reads the same:
  particle_gpu                      src/particle_gpu.h:36-99
  grid_gpu_green / grid_gpu_rothlin src/grid_gpu_base.h:82-114, grid_gpu_green.h:118-119
  tool_3d_gpu, tool_forces          src/tool_3d_gpu.h:52-176
  blanking                          src/blanking.h:29-64
  leap_frog                         src/leap_frog.h:37-50
  actions_setup_* / interactions_setup_*   src/actions_gpu.h:71-75, src/interactions_gpu.h:133-141
  global_tool, global_tool_forces, global_blanking, global_time_*   src/mfree_gpu_3d.cpp:41-48
Errors raise MfgError (the reference prints and exits, src/types.cpp:20-49).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (MFG_CONTACT_EXHAUSTIVE, MFG_CONTACT_HASHED, MFG_FP32_KAHAN, MFG_FP32_WIDE, MFG_FP64, MFG_GRID_GREEN,
                   MFG_GRID_ROTHLIN, MFG_HEAT_BROOKSHAW, MFG_HEAT_NONE, MFG_HEAT_PSE, MfgError, check,
                   corr_constants, joco_constants, phys_constants, trml_constants)

FLT_MAX = 3.4028234663852886e38

# ---- process globals (src/mfree_gpu_3d.cpp:41-48) --------------------------------------------
global_tool: list = []
global_tool_forces = None
global_blanking = None
global_time_dt = 0.0
global_time_final = 0.0
global_time_current = 0.0
global_time_step = 0


def reset_globals() -> None:
    global global_tool, global_tool_forces, global_blanking
    global global_time_dt, global_time_final, global_time_current, global_time_step
    global_tool = []
    global_tool_forces = None
    global_blanking = None
    global_time_dt = global_time_final = global_time_current = 0.0
    global_time_step = 0


def _L():
    return _lib.load()


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _arr(a, width=None) -> np.ndarray | None:
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if width is not None:
        a = a.reshape(-1, width) if a.ndim == 1 else a
    return a


# ---- constants (src/constants_structs.cpp:20-44) ---------------------------------------------
def make_phys_constants() -> phys_constants:
    return phys_constants()


def make_trml_constants() -> trml_constants:
    return trml_constants()


def make_corr_constants() -> corr_constants:
    return corr_constants()


def make_joco_constants() -> joco_constants:
    return joco_constants()


def constants_reset() -> None:
    """Clears every constant and switch of the *_setup_* calls (see mfg_constants_reset)."""
    _lib.init()
    check(_L().mfg_constants_reset())


MFG_PAIR_LIST, MFG_PAIR_TILED = 0, 1


def set_options(cspm: bool = False, heat: int = MFG_HEAT_BROOKSHAW, single_sqrt: bool = True,
                max_neighbours: int = 0, pair_pass: int = MFG_PAIR_LIST, pair_strip: int = 0,
                cell_tile: int = 0, sort_interval: int = 0, pair_block: int = 0) -> None:
    """Switches that are compile-time macros in the reference (src/types.h:32-108), the
    pair-pass kernel (list gather or cell-tiled: the same sums in the same order), the list
    pass's visiting order (strip width in grid columns; 0 default, negative the sorted order),
    the neighbour pass's cell tile (cells per wavefront: 1, 2, 4, 8; 0 automatic), the opt-in
    re-sort interval (sort every k-th step on a locked, still grid; 0 or 1 every step, as the
    reference: other values change the neighbour lists and results) and the list pair pass's
    workgroup size (64, 128, 256; 0 selects 256; the sums do not depend on it)."""
    _lib.init()
    check(_L().mfg_set_options(C.byref(_lib.options(int(cspm), int(heat), int(single_sqrt), int(max_neighbours),
                                                    int(pair_pass), int(pair_strip), int(cell_tile),
                                                    int(sort_interval), int(pair_block)))))


def actions_setup_physical_constants(c: phys_constants) -> None:
    _lib.init()
    check(_L().mfg_actions_setup_physical_constants(C.byref(c)))


def actions_setup_corrector_constants(c: corr_constants) -> None:
    _lib.init()
    check(_L().mfg_actions_setup_corrector_constants(C.byref(c)))


def actions_setup_johnson_cook_constants(c: joco_constants) -> None:
    _lib.init()
    check(_L().mfg_actions_setup_johnson_cook_constants(C.byref(c)))


def actions_setup_thermal_constants_wp(c: trml_constants) -> None:
    _lib.init()
    check(_L().mfg_actions_setup_thermal_constants_wp(C.byref(c)))


def actions_setup_thermal_constants_tool(c: trml_constants) -> None:
    _lib.init()
    check(_L().mfg_actions_setup_thermal_constants_tool(C.byref(c)))


def interactions_setup_physical_constants(c: phys_constants) -> None:
    _lib.init()
    check(_L().mfg_interactions_setup_physical_constants(C.byref(c)))


def interactions_setup_corrector_constants(c: corr_constants) -> None:
    _lib.init()
    check(_L().mfg_interactions_setup_corrector_constants(C.byref(c)))


def interactions_setup_thermal_constants_workpiece(c: trml_constants) -> None:
    _lib.init()
    check(_L().mfg_interactions_setup_thermal_constants_workpiece(C.byref(c)))


def interactions_setup_thermal_constants_tool(c: trml_constants, tool: "tool_3d_gpu | None" = None) -> None:
    _lib.init()
    check(_L().mfg_interactions_setup_thermal_constants_tool(C.byref(c), tool._h if tool else None))


def interactions_setup_geometry_constants(grid: "grid_base") -> None:
    check(_L().mfg_interactions_setup_geometry_constants(grid._h))


# ---- particle_gpu ----------------------------------------------------------------------------
class particle_gpu:
    """Particle state on the device (src/particle_gpu.h:36-99).

    Arrays are host arrays: pos / vel as (N, 3) or (N, 4), S as (N, 9) in GLM column-major order.
    """

    def __init__(self, pos, vel=None, rho=None, T_init=None, h=None, fixed=None, tool_p=None, S=None,
                 precision: int = MFG_FP64):
        _lib.init()
        pos = _arr(pos)
        n = pos.shape[0]

        def four(a):
            if a is None:
                return None
            a = _arr(a)
            if a.shape[1] == 4:
                return np.ascontiguousarray(a)
            out = np.zeros((n, 4))
            out[:, :3] = a
            return out

        arrays = [four(pos), four(vel), _arr(rho), _arr(T_init), _arr(h), _arr(fixed), _arr(tool_p), _arr(S)]
        self._keep = arrays
        ptrs = [None if a is None else _dptr(a) for a in arrays]
        handle = C.c_void_p()
        check(_L().mfg_particle_gpu_create(n, *ptrs, int(precision), C.byref(handle)))
        self._h = handle
        self.N_init = n
        self.precision = precision

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _L is not None:  # at interpreter exit the module may already be torn down
            _L().mfg_particle_gpu_destroy(h)
            self._h = None

    @property
    def N(self) -> int:
        n, ni = C.c_uint(0), C.c_uint(0)
        check(_L().mfg_particle_gpu_counts(self._h, C.byref(n), C.byref(ni)))
        return n.value

    def counts(self):
        n, ni = C.c_uint(0), C.c_uint(0)
        check(_L().mfg_particle_gpu_counts(self._h, C.byref(n), C.byref(ni)))
        return n.value, ni.value

    def pair_stats(self):
        """(candidate pairs, interacting pairs) of the last neighbour pass, which must have
        counted them (pair_stats_enable)."""
        a, b = C.c_ulonglong(0), C.c_ulonglong(0)
        check(_L().mfg_particle_gpu_pair_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def download(self, field: int) -> np.ndarray:
        w = _L().mfg_field_width(field)
        if w <= 0:
            raise MfgError(_lib.MFG_ERR_ARG, f"unknown field {field}")
        out = np.zeros(self.N_init * w)
        check(_L().mfg_particle_gpu_download(self._h, field, _dptr(out)))
        return out.reshape(self.N_init, w) if w > 1 else out

    def upload(self, field: int, values) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        check(_L().mfg_particle_gpu_upload(self._h, field, _dptr(v)))

    def snapshot(self) -> np.ndarray:
        """The persistent state as of the last queued step (bytes; see mfg_particle_gpu_snapshot)."""
        n = C.c_ulonglong(0)
        check(_L().mfg_particle_gpu_snapshot_bytes(self._h, C.byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        check(_L().mfg_particle_gpu_snapshot(self._h, buf.ctypes.data_as(C.c_void_p)))
        return buf

    def restore(self, snap: np.ndarray) -> None:
        buf = np.ascontiguousarray(snap, dtype=np.uint8)
        check(_L().mfg_particle_gpu_restore(self._h, buf.ctypes.data_as(C.c_void_p)))


# ---- grid ----------------------------------------------------------------------------------
class grid_base:
    """Spatial hashing grid (src/grid_gpu_base.h:82-114)."""

    kind = MFG_GRID_GREEN

    def __init__(self, *args):
        _lib.init()
        handle = C.c_void_p()
        if len(args) == 4:  # (num_part, bbmin, bbmax, h): locked geometry
            num_part, bbmin, bbmax, h = args
            a, b = _arr(bbmin), _arr(bbmax)
            check(_L().mfg_grid_create_locked(self.kind, int(num_part), _dptr(a), _dptr(b), float(h),
                                              C.byref(handle)))
        elif len(args) == 2:  # (max_cell, num_part): geometry follows the particles
            max_cell, num_part = args
            check(_L().mfg_grid_create_unlocked(self.kind, int(max_cell), int(num_part), C.byref(handle)))
        else:
            raise TypeError("grid(num_part, bbmin, bbmax, h) or grid(max_cell, num_part)")
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _L is not None:  # at interpreter exit the module may already be torn down
            _L().mfg_grid_destroy(h)
            self._h = None

    def _geom(self):
        g = _lib.geom_constants()
        nc = C.c_long(0)
        locked = C.c_int(0)
        check(_L().mfg_grid_get_geometry(self._h, C.byref(g), C.byref(nc), C.byref(locked)))
        return g, nc.value, bool(locked.value)

    def nx(self): return self._geom()[0].nx
    def ny(self): return self._geom()[0].ny
    def nz(self): return self._geom()[0].nz
    def bbmin_x(self): return self._geom()[0].bbmin_x
    def bbmin_y(self): return self._geom()[0].bbmin_y
    def bbmin_z(self): return self._geom()[0].bbmin_z
    def bbmax_x(self): return self._geom()[0].bbmax_x
    def bbmax_y(self): return self._geom()[0].bbmax_y
    def bbmax_z(self): return self._geom()[0].bbmax_z
    def dx(self): return self._geom()[0].dx
    def is_locked(self): return self._geom()[2]
    def num_cell(self): return self._geom()[1]

    def set_bbox_vel(self, v):
        a = _arr(v)
        check(_L().mfg_grid_set_bbox_vel(self._h, _dptr(a)))

    def set_max_fixed(self, fixed: bool): check(_L().mfg_grid_set_max_fixed(self._h, int(fixed)))
    def set_min_fixed(self, fixed: bool): check(_L().mfg_grid_set_min_fixed(self._h, int(fixed)))
    def set_hard_blank(self, hb: bool): check(_L().mfg_grid_set_hard_blank(self._h, int(hb)))

    def update_geometry(self, particles: particle_gpu, kernel_width: float = 2.0):
        check(_L().mfg_grid_update_geometry(self._h, particles._h, float(kernel_width), global_time_current))

    def assign_hashes(self, particles: particle_gpu):
        check(_L().mfg_grid_assign_hashes(self._h, particles._h))

    def adapt_particle_number(self, particles: particle_gpu):
        check(_L().mfg_grid_adapt_particle_number(self._h, particles._h))

    def sort(self, particles: particle_gpu):
        check(_L().mfg_grid_sort(self._h, particles._h))

    def get_cells(self, particles: particle_gpu | None = None):
        """Host copies of (cell_start, cell_end) in this grid's table layout."""
        n = self.num_cell()
        s = np.zeros(n, dtype=np.int32)
        e = np.zeros(n, dtype=np.int32)
        check(_L().mfg_grid_download_cells(self._h, s.ctypes.data_as(C.POINTER(C.c_int)),
                                           e.ctypes.data_as(C.POINTER(C.c_int))))
        return s, e

    def neighbours(self, particles: particle_gpu, cap: int = 0):
        """Neighbour lists of the current sorted particles, in the reference's pair-loop order."""
        counts = np.zeros(particles.N_init, dtype=np.int32)
        lists = np.full(particles.N_init * max(cap, 1), -1, dtype=np.int32) if cap > 0 else None
        ip = C.POINTER(C.c_int)
        check(_L().mfg_grid_neighbours(self._h, particles._h, counts.ctypes.data_as(ip),
                                       lists.ctypes.data_as(ip) if lists is not None else None, int(cap)))
        n = particles.N
        if lists is None:
            return counts[:n], None
        return counts[:n], lists.reshape(particles.N_init, cap)[:n]


class grid_gpu_green(grid_base):
    kind = MFG_GRID_GREEN


class grid_gpu_rothlin(grid_base):
    kind = MFG_GRID_ROTHLIN


# ---- tool ----------------------------------------------------------------------------------
class tool_3d_gpu:
    """Rigid triangle-mesh tool (src/tool_3d_gpu.h:74-176), built from an indexed mesh."""

    exhaustive = MFG_CONTACT_EXHAUSTIVE
    spatial_hashing = MFG_CONTACT_HASHED

    def __init__(self, verts, tris, vel=None, phys: phys_constants | None = None, alpha: float = 0.1,
                 h_avg: float = 0.0):
        _lib.init()
        v = _arr(verts, 3)
        t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        vv = _arr(vel) if vel is not None else None
        handle = C.c_void_p()
        check(_L().mfg_tool_3d_gpu_create(_dptr(v), v.shape[0], t.ctypes.data_as(C.POINTER(C.c_int)), t.shape[0],
                                          float(h_avg), _dptr(vv) if vv is not None else None,
                                          C.byref(phys) if phys is not None else None, float(alpha),
                                          C.byref(handle)))
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _L is not None:  # at interpreter exit the module may already be torn down
            _L().mfg_tool_destroy(h)
            self._h = None

    def update_tool(self):
        check(_L().mfg_tool_update(self._h, global_time_current, global_time_dt))

    def compute_contact_force(self, particles: particle_gpu, record_force: bool = False) -> None:
        """tool_3d_gpu::compute_contact_force (src/tool_3d_gpu.h:120): fc, ft and the contact
        normal of every particle in contact; with record_force the tool's force goes into its
        slot (its index in global_tool) of global_tool_forces."""
        forces, slot = None, 0
        if record_force and global_tool_forces is not None:
            forces = global_tool_forces._h
            slot = next(k for k, t in enumerate(global_tool) if t is self)
        check(_L().mfg_tool_compute_contact_force(self._h, particles._h, int(forces is not None), forces, slot,
                                                  global_time_dt, global_time_current))

    def set_vel(self, v):
        a = _arr(v)
        check(_L().mfg_tool_set_vel(self._h, _dptr(a)))

    def set_mu(self, mu): check(_L().mfg_tool_set_mu(self._h, float(mu)))
    def set_contact_alpha(self, a): check(_L().mfg_tool_set_contact_alpha(self._h, float(a)))
    def set_algorithm_type(self, algo): check(_L().mfg_tool_set_algorithm_type(self._h, int(algo)))
    def set_birth_death(self, b, d): check(_L().mfg_tool_set_birth_death(self._h, float(b), float(d)))
    def set_thermal(self, th): check(_L().mfg_tool_set_thermal(self._h, int(th)))
    def set_physics(self, ph: phys_constants): check(_L().mfg_tool_set_physics(self._h, C.byref(ph)))

    def get_vel(self):
        v = np.zeros(3)
        check(_L().mfg_tool_get_vel(self._h, _dptr(v)))
        return v

    def get_bbox(self):
        a, b = np.zeros(3), np.zeros(3)
        check(_L().mfg_tool_get_bbox(self._h, _dptr(a), _dptr(b)))
        return a, b

    def get_motion(self) -> np.ndarray:
        """(time the triangles were last moved to, bbmin[3], bbmax[3]) for snapshots."""
        m = np.zeros(7)
        check(_L().mfg_tool_get_motion(self._h, _dptr(m)))
        return m

    def set_motion(self, m) -> None:
        v = np.ascontiguousarray(m, dtype=np.float64).reshape(7)
        check(_L().mfg_tool_set_motion(self._h, _dptr(v)))

    def mu(self):
        m = C.c_double(0.)
        check(_L().mfg_tool_get_mu(self._h, C.byref(m)))
        return m.value

    def is_alive(self):
        return bool(check(_L().mfg_tool_is_alive(self._h, global_time_current)))

    def grid_info(self):
        dims = np.zeros(5, dtype=np.int64)
        bb = np.zeros(7)
        check(_L().mfg_tool_get_grid(self._h, dims.ctypes.data_as(C.POINTER(C.c_long)), _dptr(bb)))
        nc = int(dims[0] * dims[1] * dims[2])
        cells = np.zeros(nc + 1, dtype=np.uint32)
        boxed = np.zeros((int(dims[3]), 30))
        check(_L().mfg_tool_download_cells(self._h, cells.ctypes.data_as(C.POINTER(C.c_uint)), _dptr(boxed)))
        return dict(dims=dims[:3].copy(), n_boxed=int(dims[3]), n_tris=int(dims[4]), bbmin=bb[:3].copy(),
                    bbmax=bb[3:6].copy(), dx=float(bb[6]), cells=cells, boxed=boxed)


class tool_forces:
    """Per-tool force sums (src/tool_3d_gpu.h:52-68)."""

    def __init__(self, num_tool: int):
        _lib.init()
        handle = C.c_void_p()
        check(_L().mfg_tool_forces_create(int(num_tool), C.byref(handle)))
        self._h = handle
        self.num_tool = int(num_tool)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _L is not None:  # at interpreter exit the module may already be torn down
            _L().mfg_tool_forces_destroy(h)
            self._h = None

    def reset(self):
        check(_L().mfg_tool_forces_reset(self._h))

    def get_forces_host(self) -> np.ndarray:
        out = np.zeros(3 * self.num_tool, dtype=np.float32)
        check(_L().mfg_tool_forces_get_host(self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out.reshape(-1, 3)

    def report(self, fp) -> None:
        """One line of "%.9g %.9g %.9g " per tool, as tool_forces::report writes it (:53-69)."""
        for f in self.get_forces_host():
            fp.write("%.9g %.9g %.9g " % (f[0], f[1], f[2]))
        fp.write("\n")


# ---- blanking (host logic, src/blanking.cpp:20-67) -------------------------------------------
class blanking:
    def __init__(self, bbmin=None, bbmax=None, bbvel=None, max_vel=None):
        self.m_bbmin = np.zeros(3) if bbmin is None else np.array(bbmin, dtype=np.float64)
        self.m_bbmax = np.zeros(3) if bbmax is None else np.array(bbmax, dtype=np.float64)
        self.m_bbmin_init = self.m_bbmin.copy()
        self.m_bbmax_init = self.m_bbmax.copy()
        self.m_vel_bbox = np.zeros(3) if bbvel is None else np.array(bbvel, dtype=np.float64)
        if max_vel is None:
            self.m_max_vel_squared = FLT_MAX
        elif bbmin is None:
            # blanking(float_t max_vel) stores the value unsquared
            self.m_max_vel_squared = float(max_vel)
        else:
            self.m_max_vel_squared = float(max_vel) * float(max_vel)
        self.m_min_fixed = False
        self.m_max_fixed = False

    # the reference swaps these two (src/blanking.cpp:35-41)
    def set_max_fixed(self, fixed: bool): self.m_min_fixed = bool(fixed)
    def set_min_fixed(self, fixed: bool): self.m_max_fixed = bool(fixed)

    def update(self):
        if not self.m_min_fixed:
            self.m_bbmin = self.m_bbmin_init + global_time_current * self.m_vel_bbox
        if not self.m_max_fixed:
            self.m_bbmax = self.m_bbmax_init + global_time_current * self.m_vel_bbox

    def is_static(self) -> bool:
        if np.array_equal(self.m_bbmin, self.m_bbmax):
            return True
        return float(np.linalg.norm(self.m_vel_bbox)) == 0.0

    def get_bb(self):
        return self.m_bbmin.copy(), self.m_bbmax.copy()

    def get_max_vel_squared(self) -> float:
        return self.m_max_vel_squared

    def state(self) -> _lib.blanking_state:
        b = _lib.blanking_state()
        for k in range(3):
            b.bbmin[k] = self.m_bbmin[k]
            b.bbmax[k] = self.m_bbmax[k]
        b.max_vel_squared = self.m_max_vel_squared
        return b


def perform_blanking(particles: particle_gpu, bl: blanking) -> None:
    st = bl.state()
    check(_L().mfg_perform_blanking(particles._h, C.byref(st)))


def debug_check_valid_full(particles: particle_gpu) -> None:
    check(_L().mfg_debug_check_valid(particles._h))


# ---- the reference's actions and pair passes, one call each (src/actions_gpu.h:41-80,
# src/interactions_gpu.h:51-52); time steps from the process globals, as the reference reads them
def material_eos(particles: particle_gpu) -> None:
    check(_L().mfg_actions_material_eos(particles._h))


def corrector_artificial_stress(particles: particle_gpu) -> None:
    check(_L().mfg_actions_corrector_artificial_stress(particles._h))


def material_stress_rate_jaumann(particles: particle_gpu) -> None:
    check(_L().mfg_actions_material_stress_rate_jaumann(particles._h))


def material_fric_heat_gen(particles: particle_gpu, vel) -> None:
    v = np.ascontiguousarray(vel, dtype=np.float64).reshape(3)
    check(_L().mfg_actions_material_fric_heat_gen(particles._h, _dptr(v), global_time_dt))


def contmech_continuity(particles: particle_gpu) -> None:
    check(_L().mfg_actions_contmech_continuity(particles._h))


def contmech_momentum(particles: particle_gpu) -> None:
    check(_L().mfg_actions_contmech_momentum(particles._h))


def contmech_advection(particles: particle_gpu) -> None:
    check(_L().mfg_actions_contmech_advection(particles._h))


def plasticity_johnson_cook(particles: particle_gpu) -> None:
    check(_L().mfg_actions_plasticity_johnson_cook(particles._h, global_time_dt))


def perform_boundary_conditions_thermal(particles: particle_gpu) -> None:
    check(_L().mfg_actions_perform_boundary_conditions_thermal(particles._h))


def perform_boundary_conditions(particles: particle_gpu) -> None:
    check(_L().mfg_actions_perform_boundary_conditions(particles._h))


def actions_move_tool_particles(particles: particle_gpu, tool: "tool_3d_gpu") -> None:
    check(_L().mfg_actions_move_tool_particles(particles._h, tool._h, global_time_dt))


def debug_invalidate(particles: particle_gpu) -> None:
    check(_L().mfg_debug_invalidate(particles._h))


def interactions_monaghan(particles: particle_gpu, g: "grid_base") -> None:
    """The pair pass on the current positions with g's cell tables of the current order (the
    reference takes g's cell_start / cell_end)."""
    check(_L().mfg_interactions_monaghan(particles._h, g._h))


def interactions_heat_pse(particles: particle_gpu, g: "grid_base") -> None:
    check(_L().mfg_interactions_heat_pse(particles._h, g._h))


# ---- leap frog (src/leap_frog.h:37-50) --------------------------------------------------------
class leap_frog:
    def __init__(self, num_part: int, num_cell: int):
        _lib.init()
        handle = C.c_void_p()
        check(_L().mfg_leap_frog_create(int(num_part), int(num_cell), C.byref(handle)))
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _L is not None:  # at interpreter exit the module may already be torn down
            _L().mfg_leap_frog_destroy(h)
            self._h = None

    # leap_frog's phases (private in the reference, src/leap_frog.cu:30-119), for stepping call by call
    def init(self, particles: particle_gpu) -> None:
        check(_L().mfg_leap_frog_init(self._h, particles._h))

    def predict(self, particles: particle_gpu) -> None:
        check(_L().mfg_leap_frog_advance(self._h, particles._h, 0.5 * global_time_dt))

    def correct(self, particles: particle_gpu) -> None:
        check(_L().mfg_leap_frog_advance(self._h, particles._h, global_time_dt))

    def reset_contact_forces(self, particles: particle_gpu) -> None:
        check(_L().mfg_leap_frog_reset_contact_forces(self._h, particles._h))

    def step_by_actions(self, particles: particle_gpu, g: grid_base, record_forces: bool = False) -> None:
        """leap_frog::step (src/leap_frog.cu:121-195) composed call by call from the per-function
        entry points, in the reference's order (the invalidation always on, as in step).  The same
        results as step(), one launch per call."""
        if global_blanking is not None:
            global_blanking.update()
            perform_blanking(particles, global_blanking)
        g.update_geometry(particles)
        g.assign_hashes(particles)
        if global_blanking is not None:
            g.adapt_particle_number(particles)
        g.sort(particles)
        self.init(particles)
        self.predict(particles)
        material_eos(particles)
        corrector_artificial_stress(particles)
        interactions_monaghan(particles, g)
        opts = _lib.options()
        check(_L().mfg_get_options(C.byref(opts)))
        if opts.heat == MFG_HEAT_PSE:
            interactions_heat_pse(particles, g)
        material_stress_rate_jaumann(particles)
        contmech_continuity(particles)
        contmech_momentum(particles)
        contmech_advection(particles)
        debug_invalidate(particles)
        self.correct(particles)
        plasticity_johnson_cook(particles)
        self.reset_contact_forces(particles)
        if record_forces and global_tool_forces is not None:
            global_tool_forces.reset()
        for tool in global_tool:
            tool.update_tool()
            tool.compute_contact_force(particles, record_forces)
        if global_tool:
            material_fric_heat_gen(particles, global_tool[0].get_vel())
        perform_boundary_conditions(particles)
        perform_boundary_conditions_thermal(particles)
        if global_tool:
            actions_move_tool_particles(particles, global_tool[0])

    def step(self, particles: particle_gpu, g: grid_base, record_forces: bool = False,
             allocation_free: bool = False) -> None:
        """One time step with the process globals, as leap_frog::step (src/leap_frog.cu:121-195).
        `allocation_free`: through mfg_leap_frog_step_noalloc (same results, no device allocation;
        its work memory is the library's fixed per-GPU array)."""
        tools = global_tool
        arr = (C.c_void_p * max(len(tools), 1))(*[t._h for t in tools])
        bstate = None
        if global_blanking is not None:
            global_blanking.update()
            bstate = global_blanking.state()
        forces = global_tool_forces._h if (record_forces and global_tool_forces is not None) else None
        fn = _L().mfg_leap_frog_step_noalloc if allocation_free else _L().mfg_leap_frog_step
        check(fn(self._h, particles._h, g._h, int(bool(record_forces and forces)), arr, len(tools),
                 C.byref(bstate) if bstate is not None else None, forces, global_time_dt, global_time_current))


# ---- batched steps (no counterpart in the reference) -------------------------------------------
class batch:
    """Many independent instances stepped together: every kernel of the step runs once over all
    of them, each instance with its own grid, constants, tools and clock (mfg_batch_*).  An
    instance takes the library's constants as they are when it is added, so set them up for the
    instance (as a reference set-up does) right before `add`.  Results per instance are those of
    leap_frog.step on it alone."""

    def __init__(self, cell_tile: int = 0):
        _lib.init()
        handle = C.c_void_p()
        check(_L().mfg_batch_create(C.byref(handle)))
        self._h = handle
        self._members = []  # the objects each instance uses stay alive while it is in the batch
        if cell_tile:
            self.set_cell_tile(cell_tile)

    def set_cell_tile(self, cell_tile: int) -> None:
        """Cells per wavefront of the neighbour pass (1, 2, 4, 8; 0 automatic)."""
        check(_L().mfg_batch_set_cell_tile(self._h, int(cell_tile)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _L is not None:
            _L().mfg_batch_destroy(h)
            self._h = None

    def __len__(self) -> int:
        return len(self._members)

    def add(self, particles: particle_gpu, g: grid_base, tools=(), bl: "blanking | None" = None,
            forces: "tool_forces | None" = None, dt: float | None = None, t_current: float | None = None) -> int:
        """Adds an instance (defaults: the process globals' time step and clock); returns its slot."""
        tools = list(tools)
        arr = (C.c_void_p * max(len(tools), 1))(*[t._h for t in tools])
        bstate = None
        if bl is not None:
            if not bl.is_static():
                raise ValueError("a batched instance needs a blanking box that does not move")
            bl.update()
            bstate = bl.state()
        slot = C.c_int(-1)
        check(_L().mfg_batch_add(self._h, particles._h, g._h, arr, len(tools),
                                 C.byref(bstate) if bstate is not None else None,
                                 forces._h if forces is not None else None,
                                 global_time_dt if dt is None else float(dt),
                                 global_time_current if t_current is None else float(t_current), C.byref(slot)))
        self._members.append((particles, g, tools, bl, forces))
        return slot.value

    def remove(self, slot: int) -> None:
        check(_L().mfg_batch_remove(self._h, int(slot)))
        del self._members[slot]

    def time(self, slot: int) -> float:
        """The instance's clock (global_time_current before its next step)."""
        t = C.c_double(0.0)
        check(_L().mfg_batch_get_time(self._h, int(slot), C.byref(t)))
        return t.value

    def step(self, n_steps: int = 1, record_forces: bool = False, use_graph: bool = False) -> None:
        """Queues n_steps steps of every instance on the current stream; use_graph: one captured
        step replayed n_steps times."""
        check(_L().mfg_batch_step(self._h, int(n_steps), int(bool(record_forces)), int(bool(use_graph))))


# ---- profiling ---------------------------------------------------------------------------------
def pair_stats_enable(on: bool = True) -> None:
    """Whether later neighbour passes count their candidate and interacting pairs (read with
    particle_gpu.pair_stats); off by default."""
    check(_L().mfg_pair_stats_enable(int(bool(on))))


def profile_enable(on: int = True) -> None:
    """0 / False: off; True / MFG_PROFILE_ALL: events around every kernel; MFG_PROFILE_PASSES:
    around the neighbour and pair passes only."""
    check(_L().mfg_profile_enable(int(on)))


def profile_reset() -> None:
    check(_L().mfg_profile_reset())


def profile_read() -> dict:
    """{kernel: (launches, total_ms)} of the events recorded since the last reset."""
    out = {}
    n = _L().mfg_profile_count()
    buf = C.create_string_buffer(48)
    for k in range(n):
        launches = C.c_long(0)
        ms = C.c_double(0.)
        check(_L().mfg_profile_get(k, buf, C.byref(launches), C.byref(ms)))
        out[buf.value.decode()] = (launches.value, ms.value)
    return out


def synchronize() -> None:
    check(_L().mfg_synchronize())


def stream_pool(n: int) -> None:
    """n streams for independent instances (stream 0 is the library stream)."""
    check(_L().mfg_stream_pool(n))


def select_stream(k: int) -> None:
    """Stream k of the pool takes every later call (and synchronize waits for it)."""
    check(_L().mfg_select_stream(k))


def device_synchronize() -> None:
    """Waits for every stream of the pool."""
    check(_L().mfg_device_synchronize())


# ---- tool mesh readers ---------------------------------------------------------------------
def _read_mesh(call, *args):
    nv, nt = C.c_int(0), C.c_int(0)
    rc = call(*args, C.byref(nv), C.byref(nt), None, None)
    if rc != 0:
        raise MfgError(rc, _L().mfg_mesh_last_error().decode())
    verts = np.zeros((nv.value, 3))
    tris = np.zeros((nt.value, 3), dtype=np.int32)
    rc = call(*args, C.byref(nv), C.byref(nt), verts.ctypes.data_as(C.POINTER(C.c_double)),
              tris.ctypes.data_as(C.POINTER(C.c_int)))
    if rc != 0:
        raise MfgError(rc, _L().mfg_mesh_last_error().decode())
    return verts, tris


def read_mesh_ldynak(path: str, quads: bool = False):
    """(verts (n, 3), tris (m, 3)) of the surface of an LS-DYNA keyword tet (or hex) mesh,
    ldynak_read_triangles_from_tetmesh / _quadmesh (src/ldynak_reader.cpp:308-434)."""
    return _read_mesh(_L().mfg_mesh_read_ldynak, path.encode(), int(bool(quads)))


def read_mesh_json(path: str):
    """(verts, tris) of a MeshLab JSON triangle mesh, json_read_triangles (src/json_reader.cpp:108-117)."""
    return _read_mesh(_L().mfg_mesh_read_json, path.encode())


# ---- warm starts ----------------------------------------------------------------------------
def _read_fields(call, path: str, width: int):
    n = C.c_int(0)
    rc = call(path.encode(), C.byref(n), None)
    if rc != 0:
        raise MfgError(rc, _L().mfg_mesh_last_error().decode())
    out = np.zeros((n.value, width))
    rc = call(path.encode(), C.byref(n), out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise MfgError(rc, _L().mfg_mesh_last_error().decode())
    return out


def read_vtk_positions(path: str) -> np.ndarray:
    """(n, 3) positions of a legacy VTK particle file, vtk_read_init_pos (src/vtk_reader.cpp:29-69)."""
    return _read_fields(_L().mfg_vtk_read_positions, path, 3)


def read_steady_temperatures(path: str) -> np.ndarray:
    """(n, 4) steady-state samples x y z T, as the single-grain set-up reads them
    (src/Benchmarks/benchmarks_single_grain.cu:281-301)."""
    return _read_fields(_L().mfg_read_steady_temperatures, path, 4)


def interp_temps(samples: np.ndarray, weights: float, h: float, points: np.ndarray) -> np.ndarray:
    """interp_temps (src/Benchmarks/interp_utils.cpp:65-99) on the GPU: for every point,
    sum_j weights * W(|p - x_j|, h) * v_j over the samples (x, y, z, v) in order."""
    _lib.init()
    a = np.ascontiguousarray(samples, dtype=np.float64).reshape(-1, 4)
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(p.shape[0])
    dp = C.POINTER(C.c_double)
    check(_L().mfg_interp_temps(a.ctypes.data_as(dp), a.shape[0], float(weights), float(h),
                                p.ctypes.data_as(dp), p.shape[0], out.ctypes.data_as(dp)))
    return out
