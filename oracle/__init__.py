"""CPU FP64 oracle of the reference SPH cutting step -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
this module, always as the checker of the HIP product path.  The product library never loads it.

The C++ restatement lives in ``mfree_oracle.cpp`` (each function cites the reference file:line it
follows); this module is a thin ctypes wrapper with the same entry points.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SRC = os.path.join(_HERE, "mfree_oracle.cpp")
_LIB = os.path.join(_HERE, "liboracle.so")

# field ids shared with include/mfree_gpu.h (MFG_F_*)
F_POS, F_VEL, F_VEL_BC, F_FC, F_FT, F_H, F_RHO, F_P, F_S, F_R = range(10)
F_EPS_PL, F_EPS_PL_DOT, F_T, F_FIXED, F_BLANKED, F_TOOL = range(10, 16)
F_POS_T, F_VEL_T, F_RHO_T, F_S_T, F_T_T, F_S_DER, F_V_DER = range(16, 23)
F_HASH, F_IDX, F_ID, F_N = 23, 24, 25, 26

HEAT_NONE, HEAT_BROOKSHAW, HEAT_PSE = 0, 1, 2


def build(force: bool = False) -> str:
    """Compile the oracle with g++ (-O2, OpenMP, no FP contraction)."""
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(_SRC):
        cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
               "-o", _LIB + ".tmp", _SRC]
        subprocess.run(cmd, check=True)
        os.replace(_LIB + ".tmp", _LIB)
    return _LIB


def build_native() -> str:
    """The CPU-baseline build of the same source: g++ -O3 -march=native with OpenMP, still without
    FP contraction (so its results are the -O2 build's), compiled for the host it runs on into
    oracle/_native/, one library per host CPU model."""
    import hashlib
    import platform
    model = platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    tag = hashlib.sha1(model.encode()).hexdigest()[:10]
    out_dir = os.path.join(_HERE, "_native")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, f"liboracle_native_{tag}.so")
    if not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(_SRC):
        cmd = ["g++", "-O3", "-march=native", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fPIC",
               "-shared", "-o", out + ".tmp", _SRC]
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)
    return out


_lib = None


def lib():
    """The oracle library: the -O2 build, or the library named by MFG_ORACLE_LIB (the CPU
    baseline's native build)."""
    global _lib
    if _lib is None:
        path = os.environ.get("MFG_ORACLE_LIB")
        if not path:
            build()
            path = _LIB
        _lib = C.CDLL(path)
        _declare(_lib)
    return _lib


class Phys(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("E", "nu", "rho0", "K", "G", "mass")]


class Trml(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("cp", "tq", "k", "alpha", "T_init", "eta")]


class Corr(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("wdeltap", "stresseps", "xspheps", "alpha", "beta", "eta")]


class Joco(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("A", "B", "C", "n", "m", "Tmelt", "Tref", "eps_dot_ref",
                                          "clamp_temp")]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_uint)
_lp = C.POINTER(C.c_long)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p


def _declare(L):
    sig = {
        "o_create": (_vp, [C.c_uint, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "o_destroy": (None, [_vp]),
        "o_set_options": (None, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "o_setup_physical_constants": (None, [_vp, C.POINTER(Phys), C.c_int]),
        "o_setup_corrector_constants": (None, [_vp, C.POINTER(Corr), C.c_int]),
        "o_setup_johnson_cook_constants": (None, [_vp, C.POINTER(Joco)]),
        "o_setup_thermal_constants_wp": (None, [_vp, C.POINTER(Trml), C.c_int]),
        "o_setup_thermal_constants_tool": (None, [_vp, C.POINTER(Trml), C.c_int]),
        "o_set_grid_locked": (None, [_vp, _dp, _dp, C.c_double]),
        "o_set_grid_unlocked": (None, [_vp, C.c_long]),
        "o_set_grid_params": (None, [_vp, _dp, C.c_int, C.c_int, C.c_int]),
        "o_get_grid": (None, [_vp, _lp, _dp]),
        "o_set_blanking": (None, [_vp, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int, C.c_int]),
        "o_add_tool": (C.c_int, [_vp, _dp, C.c_int, _ip, C.c_int, C.c_double, _dp, C.c_double,
                                 C.c_double, C.c_int, C.c_double, C.c_double, C.c_double]),
        "o_tool_set_vel_constant": (None, [_vp, C.c_int, _dp]),
        "o_tool_info": (None, [_vp, C.c_int, _lp, _dp]),
        "o_tool_cells": (None, [_vp, C.c_int, _up]),
        "o_tool_tris": (None, [_vp, C.c_int, C.c_int, _dp]),
        "o_tool_inside": (None, [_vp, C.c_int, C.c_int, _dp, _ip, _ip]),
        "o_tool_closest": (None, [_vp, C.c_int, C.c_int, _dp, _dp, _dp]),
        "o_set_time": (None, [_vp, C.c_double, C.c_double]),
        "o_set_threads": (None, [_vp, C.c_int]),
        "o_step": (C.c_int, [_vp, C.c_int]),
        "o_advance_time": (None, [_vp]),
        "o_run": (C.c_int, [_vp, C.c_int, C.c_int]),
        "o_sort_only": (C.c_int, [_vp]),
        "o_plasticity_only": (None, [_vp]),
        "o_counts": (None, [_vp, _up, _up]),
        "o_secant_failures": (C.c_long, [_vp]),
        "o_field_width": (C.c_int, [C.c_int]),
        "o_get_field": (None, [_vp, C.c_int, _dp]),
        "o_set_field": (None, [_vp, C.c_int, _dp]),
        "o_get_cells": (None, [_vp, _ip, _ip]),
        "o_get_forces": (None, [_vp, _fp]),
        "o_neighbours": (None, [_vp, _ip, _ip, C.c_int]),
        "o_cubic_spline": (None, [_dp, _dp, C.c_double, C.c_int, _dp]),
        "o_interp_temps": (None, [_dp, C.c_int, C.c_double, C.c_double, _dp, C.c_int, _dp]),
        "o_lapl_pse": (C.c_double, [_dp, _dp, C.c_double]),
        "o_sigma_yield": (C.c_double, [C.POINTER(Joco), C.c_double, C.c_double, C.c_double]),
        "o_opt_fun": (C.c_double, [C.POINTER(Joco), C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_double]),
        "o_solve_secant": (C.c_double, [C.POINTER(Joco), C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_double, _ip, _ip]),
        "o_solve_eigen": (None, [_dp, _dp, _dp, _dp, _dp]),
        "o_mat3_det_inv": (None, [_dp, _dp, _dp]),
        "o_mat3_mul": (None, [_dp, _dp, _dp]),
        "o_art_stress": (None, [C.c_double, C.c_double, C.c_double, _dp, _dp]),
        "o_jaumann": (None, [C.c_double, _dp, _dp, _dp]),
        "o_closest_point": (None, [_dp, _dp, _dp, _dp]),
        "o_ray_intersect": (C.c_int, [_dp, _dp, _dp, _dp]),
        "o_tri_aabb": (C.c_int, [_dp, _dp, _dp]),
        "o_mesh_normals": (None, [_dp, C.c_int, _ip, C.c_int, _dp]),
        "o_mesh_average_edge_length": (C.c_double, [_dp, _ip, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _d(a):
    """numpy float64 array -> ctypes pointer (None passes NULL)."""
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.ctypes.data_as(_dp)


def _keep(*arrs):
    return [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in arrs]


# ---- stand-alone functions --------------------------------------------------------------------
def cubic_spline(pi, pj, h, single_sqrt=True):
    out = np.zeros(4)
    a, b = _keep(pi, pj)
    lib().o_cubic_spline(_d(a), _d(b), h, int(single_sqrt), _d(out))
    return out


def interp_temps(samples, weights, h, points):
    """interp_temps (src/Benchmarks/interp_utils.cpp:65-99): per point, the sum over the samples
    (x, y, z, v) in order of weights * W(|p - x_j|, h) * v_j."""
    a = np.ascontiguousarray(samples, dtype=np.float64).reshape(-1, 4)
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(p.shape[0])
    lib().o_interp_temps(_d(a), a.shape[0], float(weights), float(h), _d(p), p.shape[0], _d(out))
    return out


def lapl_pse(pi, pj, h):
    a, b = _keep(pi, pj)
    return lib().o_lapl_pse(_d(a), _d(b), h)


def sigma_yield(jc: Joco, eps_pl, eps_pl_dot, t):
    return lib().o_sigma_yield(C.byref(jc), eps_pl, eps_pl_dot, t)


def opt_fun(jc: Joco, dl, norm, eps_pl, t, dt, mu):
    return lib().o_opt_fun(C.byref(jc), dl, norm, eps_pl, t, dt, mu)


def solve_secant(jc: Joco, init, tol, norm, eps_pl, t, dt, mu):
    failed = C.c_int(0)
    iters = C.c_int(0)
    r = lib().o_solve_secant(C.byref(jc), init, tol, norm, eps_pl, t, dt, mu, C.byref(failed),
                             C.byref(iters))
    return r, bool(failed.value), iters.value


def solve_eigen(a00, a01, a02, a11, a12, a22):
    a = np.array([a00, a01, a02, a11, a12, a22], dtype=np.float64)
    ev, e1, e2, e3 = np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3)
    lib().o_solve_eigen(_d(a), _d(ev), _d(e1), _d(e2), _d(e3))
    return ev, e1, e2, e3


def glm_from_rowmajor(M):
    """3x3 row-major (math) matrix -> 9 doubles in GLM storage m[col][row]."""
    return np.ascontiguousarray(np.asarray(M, dtype=np.float64).T.reshape(9))


def rowmajor_from_glm(m9):
    return np.asarray(m9, dtype=np.float64).reshape(3, 3).T.copy()


def mat3_det_inv(M):
    m9 = glm_from_rowmajor(M)
    det = np.zeros(1)
    inv = np.zeros(9)
    lib().o_mat3_det_inv(_d(m9), _d(det), _d(inv))
    return det[0], rowmajor_from_glm(inv)


def mat3_mul(A, B):
    a, b = glm_from_rowmajor(A), glm_from_rowmajor(B)
    out = np.zeros(9)
    lib().o_mat3_mul(_d(a), _d(b), _d(out))
    return rowmajor_from_glm(out)


def art_stress(eps, rho, p, S):
    s9 = glm_from_rowmajor(S)
    out = np.zeros(9)
    lib().o_art_stress(eps, rho, p, _d(s9), _d(out))
    return rowmajor_from_glm(out)


def jaumann(G, vd_glm9, S_glm9):
    """Stress rate from v_der / S given in GLM storage (v_der[a][b] = dv_a/dx_b)."""
    a, b = _keep(vd_glm9, S_glm9)
    out = np.zeros(9)
    lib().o_jaumann(G, _d(a), _d(b), _d(out))
    return out


def tri30(p1, p2, p3, n=None, np1=None, np2=None, np3=None, ne1=None, ne2=None, ne3=None):
    z = np.zeros(3)
    parts = [p1, p2, p3] + [z if v is None else v for v in (n, np1, np2, np3, ne1, ne2, ne3)]
    return np.concatenate([np.asarray(v, dtype=np.float64) for v in parts])


def closest_point(t30, qp):
    a, b = _keep(t30, qp)
    cp, nn = np.zeros(3), np.zeros(3)
    lib().o_closest_point(_d(a), _d(b), _d(cp), _d(nn))
    return cp, nn


def ray_intersect(t30, o, d):
    a, b, c = _keep(t30, o, d)
    t = np.zeros(1)
    hit = lib().o_ray_intersect(_d(a), _d(b), _d(c), _d(t))
    return bool(hit), t[0]


def tri_aabb(t30, bbmin, bbmax):
    a, b, c = _keep(t30, bbmin, bbmax)
    return bool(lib().o_tri_aabb(_d(a), _d(b), _d(c)))


def mesh_normals(verts, tris):
    v = np.ascontiguousarray(verts, dtype=np.float64)
    t = np.ascontiguousarray(tris, dtype=np.int32)
    out = np.zeros((t.shape[0], 30))
    lib().o_mesh_normals(_d(v), v.shape[0], t.ctypes.data_as(_ip), t.shape[0], _d(out))
    return out


def mesh_average_edge_length(verts, tris):
    v = np.ascontiguousarray(verts, dtype=np.float64)
    t = np.ascontiguousarray(tris, dtype=np.int32)
    return lib().o_mesh_average_edge_length(_d(v), t.ctypes.data_as(_ip), t.shape[0])


# ---- simulation -------------------------------------------------------------------------------
class OracleSim:
    """One reference simulation on the CPU: particles, locked grid, tools, blanking, time."""

    def __init__(self, pos3, vel3=None, rho=None, T=None, h=None, fixed=None, tool_p=None, S9=None):
        arrs = _keep(pos3, vel3, rho, T, h, fixed, tool_p, S9)
        self.N_init = int(np.asarray(pos3).shape[0])
        self._h = lib().o_create(self.N_init, *[_d(a) for a in arrs])
        self.n_tools = 0

    def __del__(self):
        if getattr(self, "_h", None):
            lib().o_destroy(self._h)
            self._h = None

    def set_options(self, cspm=False, heat=HEAT_BROOKSHAW, single_sqrt=True, rothlin=False):
        lib().o_set_options(self._h, int(cspm), int(heat), int(single_sqrt), int(rothlin))

    def setup_constants(self, phys: Phys, corr: Corr, joco: Joco | None, trml_wp: Trml | None,
                        trml_tool: Trml | None = None):
        L = lib()
        L.o_setup_physical_constants(self._h, C.byref(phys), 0)
        L.o_setup_corrector_constants(self._h, C.byref(corr), 0)
        if joco is not None:
            L.o_setup_johnson_cook_constants(self._h, C.byref(joco))
        if trml_wp is not None:
            L.o_setup_thermal_constants_wp(self._h, C.byref(trml_wp), 0)
        if trml_tool is not None:
            L.o_setup_thermal_constants_tool(self._h, C.byref(trml_tool), 0)

    def set_grid_locked(self, bbmin, bbmax, h):
        a, b = _keep(bbmin, bbmax)
        lib().o_set_grid_locked(self._h, _d(a), _d(b), float(h))

    def set_grid_unlocked(self, max_cell):
        """Geometry follows the particles (grid_gpu_green(max_cell, num_part))."""
        lib().o_set_grid_unlocked(self._h, int(max_cell))

    def set_grid_params(self, vel_bbox=(0, 0, 0), min_fixed=False, max_fixed=False, hard_blank=True):
        (v,) = _keep(vel_bbox)
        lib().o_set_grid_params(self._h, _d(v), int(min_fixed), int(max_fixed), int(hard_blank))

    def grid(self):
        n = np.zeros(4, dtype=np.int64)
        b = np.zeros(7)
        lib().o_get_grid(self._h, n.ctypes.data_as(_lp), _d(b))
        return dict(nx=int(n[0]), ny=int(n[1]), nz=int(n[2]), num_cell=int(n[3]), bbmin=b[:3].copy(),
                    bbmax=b[3:6].copy(), dx=float(b[6]))

    def set_blanking(self, bbmin, bbmax, vel=(0, 0, 0), max_vel_squared=3.4028234663852886e38,
                     enabled=True, min_fixed=False, max_fixed=False):
        a, b, v = _keep(bbmin, bbmax, vel)
        lib().o_set_blanking(self._h, int(enabled), _d(a), _d(b), _d(v), float(max_vel_squared),
                             int(min_fixed), int(max_fixed))

    def add_tool(self, verts, tris, vel, mass, alpha=0.1, algo=0, mu=0.0, h_avg=0.0, birth=0.0,
                 death=3.4028234663852886e38):
        v = np.ascontiguousarray(verts, dtype=np.float64)
        t = np.ascontiguousarray(tris, dtype=np.int32)
        (vv,) = _keep(vel)
        k = lib().o_add_tool(self._h, _d(v), v.shape[0], t.ctypes.data_as(_ip), t.shape[0], float(h_avg),
                             _d(vv), float(mass), float(alpha), int(algo), float(mu), float(birth),
                             float(death))
        self.n_tools = k + 1
        return k

    def tool_set_vel_constant(self, k, v):
        (vv,) = _keep(v)
        lib().o_tool_set_vel_constant(self._h, k, _d(vv))

    def tool_grid(self, k):
        dims = np.zeros(5, dtype=np.int64)
        bb = np.zeros(7)
        lib().o_tool_info(self._h, k, dims.ctypes.data_as(_lp), _d(bb))
        nc = int(dims[0] * dims[1] * dims[2])
        cells = np.zeros(nc + 1, dtype=np.uint32)
        lib().o_tool_cells(self._h, k, cells.ctypes.data_as(_up))
        boxed = np.zeros((int(dims[3]), 30))
        lib().o_tool_tris(self._h, k, 1, _d(boxed))
        tris = np.zeros((int(dims[4]), 30))
        lib().o_tool_tris(self._h, k, 0, _d(tris))
        return dict(dims=dims[:3].copy(), n_boxed=int(dims[3]), bbmin=bb[:3].copy(), bbmax=bb[3:6].copy(),
                    dx=float(bb[6]), cells=cells, boxed=boxed, tris=tris)

    def tool_inside(self, k, pts):
        p = np.ascontiguousarray(pts, dtype=np.float64)
        a = np.zeros(p.shape[0], dtype=np.int32)
        b = np.zeros(p.shape[0], dtype=np.int32)
        lib().o_tool_inside(self._h, k, p.shape[0], _d(p), a.ctypes.data_as(_ip), b.ctypes.data_as(_ip))
        return a.astype(bool), b.astype(bool)

    def tool_closest(self, k, pts):
        """Closest point and normal the hashed algorithm's shell search finds for each point."""
        p = np.ascontiguousarray(pts, dtype=np.float64)
        cp = np.zeros_like(p)
        nn = np.zeros_like(p)
        lib().o_tool_closest(self._h, k, p.shape[0], _d(p), _d(cp), _d(nn))
        return cp, nn

    def set_time(self, dt, t):
        lib().o_set_time(self._h, float(dt), float(t))

    def set_threads(self, n):
        lib().o_set_threads(self._h, int(n))

    def step(self, record=False):
        return lib().o_step(self._h, int(record))

    def advance_time(self):
        lib().o_advance_time(self._h)

    def run(self, n, record=False):
        return lib().o_run(self._h, int(n), int(record))

    def sort_only(self):
        return lib().o_sort_only(self._h)

    def plasticity_only(self):
        lib().o_plasticity_only(self._h)

    def counts(self):
        n, ni = C.c_uint(0), C.c_uint(0)
        lib().o_counts(self._h, C.byref(n), C.byref(ni))
        return n.value, ni.value

    def secant_failures(self):
        return lib().o_secant_failures(self._h)

    def field(self, f):
        w = lib().o_field_width(f)
        out = np.zeros(self.N_init * w)
        lib().o_get_field(self._h, f, _d(out))
        return out.reshape(self.N_init, w) if w > 1 else out

    def set_field(self, f, values):
        (v,) = _keep(values)
        lib().o_set_field(self._h, f, _d(v.reshape(-1)))

    def cells(self):
        g = self.grid()
        s = np.zeros(g["num_cell"], dtype=np.int32)
        e = np.zeros(g["num_cell"], dtype=np.int32)
        lib().o_get_cells(self._h, s.ctypes.data_as(_ip), e.ctypes.data_as(_ip))
        return s, e

    def forces(self):
        out = np.zeros(3 * max(self.n_tools, 1), dtype=np.float32)
        lib().o_get_forces(self._h, out.ctypes.data_as(_fp))
        return out.reshape(-1, 3)

    def neighbours(self, cap=0):
        n = self.counts()[0]
        counts = np.zeros(self.N_init, dtype=np.int32)
        if cap > 0:
            lists = np.full(self.N_init * cap, -1, dtype=np.int32)
            lib().o_neighbours(self._h, counts.ctypes.data_as(_ip), lists.ctypes.data_as(_ip), cap)
            return counts[:n], lists.reshape(self.N_init, cap)[:n]
        lib().o_neighbours(self._h, counts.ctypes.data_as(_ip), None, 0)
        return counts[:n], None
