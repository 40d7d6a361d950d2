"""ctypes bindings of the C ABI in include/mfree_gpu.h (libmfree_hip.so, built in-tree).

Loading never falls back to anything: if the HIP library is missing the import of the bindings
raises, so no GPU call can silently run elsewhere.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfree_hip.so")

# ---- constants mirrored from include/mfree_gpu.h --------------------------------------------
MFG_OK = 0
MFG_ERR_ARG, MFG_ERR_HIP, MFG_ERR_STATE, MFG_ERR_CAPACITY = -1, -2, -3, -4
MFG_ERR_ALL_BLANKED, MFG_ERR_GRID, MFG_ERR_NOMEM, MFG_ERR_NONFINITE = -5, -6, -7, -8
MFG_FP64, MFG_FP32_KAHAN, MFG_FP32_WIDE = 0, 1, 2
MFG_HEAT_NONE, MFG_HEAT_BROOKSHAW, MFG_HEAT_PSE = 0, 1, 2
MFG_CONTACT_EXHAUSTIVE, MFG_CONTACT_HASHED = 0, 1
MFG_GRID_GREEN, MFG_GRID_ROTHLIN = 0, 1
MFG_PROFILE_ALL, MFG_PROFILE_PASSES = 1, 2
MFG_TOOL_WEDGE, MFG_TOOL_GRAIN = 0, 1

F_POS, F_VEL, F_VEL_BC, F_FC, F_FT, F_H, F_RHO, F_P, F_S, F_R = range(10)
F_EPS_PL, F_EPS_PL_DOT, F_T, F_FIXED, F_BLANKED, F_TOOL = range(10, 16)
F_POS_T, F_VEL_T, F_RHO_T, F_S_T, F_T_T = range(16, 21)
F_S_DER, F_V_DER = 21, 22
F_HASH, F_IDX, F_ID, F_N = 23, 24, 25, 26


class MfgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _dfields(*names):
    return [(n, C.c_double) for n in names]


class phys_constants(C.Structure):
    _fields_ = _dfields("E", "nu", "rho0", "K", "G", "mass")


class trml_constants(C.Structure):
    _fields_ = _dfields("cp", "tq", "k", "alpha", "T_init", "eta")


class corr_constants(C.Structure):
    _fields_ = _dfields("wdeltap", "stresseps", "xspheps", "alpha", "beta", "eta")


class joco_constants(C.Structure):
    _fields_ = _dfields("A", "B", "C", "n", "m", "Tmelt", "Tref", "eps_dot_ref", "clamp_temp")


class geom_constants(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)] + _dfields(
        "bbmin_x", "bbmin_y", "bbmin_z", "bbmax_x", "bbmax_y", "bbmax_z", "dx")


class options(C.Structure):
    _fields_ = [("cspm", C.c_int), ("heat", C.c_int), ("single_sqrt", C.c_int), ("max_neighbours", C.c_int),
                ("pair_pass", C.c_int), ("pair_strip", C.c_int), ("cell_tile", C.c_int),
                ("sort_interval", C.c_int), ("pair_block", C.c_int)]


class blanking_state(C.Structure):
    _fields_ = [("bbmin", C.c_double * 3), ("bbmax", C.c_double * 3), ("max_vel_squared", C.c_double)]


class cut_config(C.Structure):
    _fields_ = ([("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)] + _dfields("dx", "jitter")
                + [("seed", C.c_ulonglong)] + _dfields("hdx", "cfl")
                + [("tool_kind", C.c_int), ("n_grains", C.c_int)]
                + _dfields("grain_offset_x", "grain_offset_y", "rake_deg", "clearance_deg", "depth_of_cut",
                           "speed", "grain_edge", "edge_radius", "facet_perturb")
                + [("grain_res", C.c_int), ("tool_particles", C.c_int)]
                + _dfields("tool_rho", "tool_cp", "tool_k", "E", "nu", "rho0", "cp", "k", "tq", "T_init",
                           "eta_fric", "jc_A", "jc_B", "jc_n", "jc_C", "jc_m", "jc_eps_dot_ref", "jc_Tref",
                           "jc_Tmelt", "jc_clamp_temp", "visc_alpha", "visc_beta", "visc_eta", "xspheps",
                           "stresseps", "mu", "contact_alpha", "max_vel")
                + [("cspm", C.c_int), ("heat", C.c_int), ("algo", C.c_int), ("steps", C.c_int)])


class scene_params(C.Structure):
    _fields_ = [("phys", phys_constants), ("corr", corr_constants), ("joco", joco_constants),
                ("trml_wp", trml_constants), ("trml_tool", trml_constants), ("dt", C.c_double),
                ("steps", C.c_int), ("tool_vel", C.c_double * 3), ("h", C.c_double),
                ("bbmin", C.c_double * 3), ("bbmax", C.c_double * 3), ("max_vel", C.c_double),
                ("cspm", C.c_int), ("heat", C.c_int), ("algo", C.c_int), ("mu", C.c_double),
                ("contact_alpha", C.c_double), ("n_tools", C.c_int), ("n_particles", C.c_uint),
                ("n_workpiece", C.c_uint)]


_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_uint)
_lp = C.POINTER(C.c_long)
_fp = C.POINTER(C.c_float)
_i = C.c_int
_d = C.c_double

# name -> (restype, argtypes); every symbol declared in include/mfree_gpu.h
SIGNATURES = {
    "mfg_init": (_i, [_i]),
    "mfg_last_error": (C.c_char_p, []),
    "mfg_synchronize": (_i, []),
    "mfg_get_stream": (_vp, []),
    "mfg_stream_pool": (_i, [_i]),
    "mfg_select_stream": (_i, [_i]),
    "mfg_device_synchronize": (_i, []),
    "mfg_set_options": (_i, [C.POINTER(options)]),
    "mfg_get_options": (_i, [C.POINTER(options)]),
    "mfg_version": (_i, []),
    "mfg_actions_setup_physical_constants": (_i, [C.POINTER(phys_constants)]),
    "mfg_actions_setup_corrector_constants": (_i, [C.POINTER(corr_constants)]),
    "mfg_actions_setup_johnson_cook_constants": (_i, [C.POINTER(joco_constants)]),
    "mfg_actions_setup_thermal_constants_wp": (_i, [C.POINTER(trml_constants)]),
    "mfg_actions_setup_thermal_constants_tool": (_i, [C.POINTER(trml_constants)]),
    "mfg_interactions_setup_physical_constants": (_i, [C.POINTER(phys_constants)]),
    "mfg_interactions_setup_corrector_constants": (_i, [C.POINTER(corr_constants)]),
    "mfg_interactions_setup_thermal_constants_workpiece": (_i, [C.POINTER(trml_constants)]),
    "mfg_interactions_setup_thermal_constants_tool": (_i, [C.POINTER(trml_constants), _vp]),
    "mfg_interactions_setup_geometry_constants": (_i, [_vp]),
    "mfg_particle_gpu_create": (_i, [C.c_uint, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i,
                                     C.POINTER(_vp)]),
    "mfg_particle_gpu_destroy": (_i, [_vp]),
    "mfg_particle_gpu_counts": (_i, [_vp, _up, _up]),
    "mfg_particle_gpu_precision": (_i, [_vp]),
    "mfg_particle_gpu_download": (_i, [_vp, _i, _dp]),
    "mfg_particle_gpu_upload": (_i, [_vp, _i, _dp]),
    "mfg_particle_gpu_snapshot_bytes": (_i, [_vp, C.POINTER(C.c_ulonglong)]),
    "mfg_particle_gpu_snapshot": (_i, [_vp, _vp]),
    "mfg_particle_gpu_restore": (_i, [_vp, _vp]),
    "mfg_field_width": (_i, [_i]),
    "mfg_particle_gpu_pair_stats": (_i, [_vp, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
    "mfg_pair_stats_enable": (_i, [_i]),
    "mfg_grid_create_locked": (_i, [_i, _i, _dp, _dp, _d, C.POINTER(_vp)]),
    "mfg_grid_create_unlocked": (_i, [_i, C.c_long, _i, C.POINTER(_vp)]),
    "mfg_grid_destroy": (_i, [_vp]),
    "mfg_grid_get_geometry": (_i, [_vp, C.POINTER(geom_constants), _lp, _ip]),
    "mfg_grid_set_bbox_vel": (_i, [_vp, _dp]),
    "mfg_grid_set_max_fixed": (_i, [_vp, _i]),
    "mfg_grid_set_min_fixed": (_i, [_vp, _i]),
    "mfg_grid_set_hard_blank": (_i, [_vp, _i]),
    "mfg_grid_update_geometry": (_i, [_vp, _vp, _d, _d]),
    "mfg_grid_assign_hashes": (_i, [_vp, _vp]),
    "mfg_grid_adapt_particle_number": (_i, [_vp, _vp]),
    "mfg_grid_sort": (_i, [_vp, _vp]),
    "mfg_grid_get_cells": (_i, [_vp, _vp, _vp, _vp]),
    "mfg_grid_download_cells": (_i, [_vp, _ip, _ip]),
    "mfg_grid_neighbours": (_i, [_vp, _vp, _ip, _ip, _i]),
    "mfg_tool_3d_gpu_create": (_i, [_dp, _i, _ip, _i, _d, _dp, C.POINTER(phys_constants), _d,
                                    C.POINTER(_vp)]),
    "mfg_tool_destroy": (_i, [_vp]),
    "mfg_tool_set_vel": (_i, [_vp, _dp]),
    "mfg_tool_set_mu": (_i, [_vp, _d]),
    "mfg_tool_set_contact_alpha": (_i, [_vp, _d]),
    "mfg_tool_set_algorithm_type": (_i, [_vp, _i]),
    "mfg_tool_set_birth_death": (_i, [_vp, _d, _d]),
    "mfg_tool_set_thermal": (_i, [_vp, _i]),
    "mfg_tool_set_physics": (_i, [_vp, C.POINTER(phys_constants)]),
    "mfg_tool_get_vel": (_i, [_vp, _dp]),
    "mfg_tool_get_bbox": (_i, [_vp, _dp, _dp]),
    "mfg_tool_get_motion": (_i, [_vp, _dp]),
    "mfg_tool_set_motion": (_i, [_vp, _dp]),
    "mfg_tool_get_mu": (_i, [_vp, _dp]),
    "mfg_tool_is_alive": (_i, [_vp, _d]),
    "mfg_tool_get_grid": (_i, [_vp, _lp, _dp]),
    "mfg_tool_download_cells": (_i, [_vp, _up, _dp]),
    "mfg_tool_update": (_i, [_vp, _d, _d]),
    "mfg_tool_forces_create": (_i, [C.c_uint, C.POINTER(_vp)]),
    "mfg_tool_forces_destroy": (_i, [_vp]),
    "mfg_tool_forces_reset": (_i, [_vp]),
    "mfg_tool_forces_get_host": (_i, [_vp, _fp]),
    "mfg_perform_blanking": (_i, [_vp, C.POINTER(blanking_state)]),
    "mfg_leap_frog_create": (_i, [C.c_uint, C.c_uint, C.POINTER(_vp)]),
    "mfg_leap_frog_destroy": (_i, [_vp]),
    "mfg_leap_frog_step": (_i, [_vp, _vp, _vp, _i, C.POINTER(_vp), _i, C.POINTER(blanking_state), _vp,
                                _d, _d]),
    "mfg_leap_frog_step_noalloc": (_i, [_vp, _vp, _vp, _i, C.POINTER(_vp), _i, C.POINTER(blanking_state),
                                        _vp, _d, _d]),
    "mfg_batch_create": (_i, [C.POINTER(_vp)]),
    "mfg_batch_destroy": (_i, [_vp]),
    "mfg_batch_add": (_i, [_vp, _vp, _vp, C.POINTER(_vp), _i, C.POINTER(blanking_state), _vp, _d, _d, _ip]),
    "mfg_batch_remove": (_i, [_vp, _i]),
    "mfg_batch_size": (_i, [_vp]),
    "mfg_batch_set_cell_tile": (_i, [_vp, _i]),
    "mfg_batch_get_time": (_i, [_vp, _i, _dp]),
    "mfg_batch_step": (_i, [_vp, _i, _i, _i]),
    "mfg_debug_check_valid": (_i, [_vp]),
    "mfg_actions_material_eos": (_i, [_vp]),
    "mfg_actions_corrector_artificial_stress": (_i, [_vp]),
    "mfg_actions_material_stress_rate_jaumann": (_i, [_vp]),
    "mfg_actions_material_fric_heat_gen": (_i, [_vp, _dp, _d]),
    "mfg_actions_contmech_continuity": (_i, [_vp]),
    "mfg_actions_contmech_momentum": (_i, [_vp]),
    "mfg_actions_contmech_advection": (_i, [_vp]),
    "mfg_actions_plasticity_johnson_cook": (_i, [_vp, _d]),
    "mfg_actions_perform_boundary_conditions_thermal": (_i, [_vp]),
    "mfg_actions_perform_boundary_conditions": (_i, [_vp]),
    "mfg_actions_move_tool_particles": (_i, [_vp, _vp, _d]),
    "mfg_debug_invalidate": (_i, [_vp]),
    "mfg_interactions_monaghan": (_i, [_vp, _vp]),
    "mfg_interactions_heat_pse": (_i, [_vp, _vp]),
    "mfg_tool_compute_contact_force": (_i, [_vp, _vp, _i, _vp, _i, _d, _d]),
    "mfg_leap_frog_init": (_i, [_vp, _vp]),
    "mfg_leap_frog_advance": (_i, [_vp, _vp, _d]),
    "mfg_leap_frog_reset_contact_forces": (_i, [_vp, _vp]),
    "mfg_profile_enable": (_i, [_i]),
    "mfg_profile_reset": (_i, []),
    "mfg_profile_count": (_i, []),
    "mfg_profile_get": (_i, [_i, C.c_char_p, _lp, _dp]),
    "mfg_scene_preset": (_i, [_i, _d, C.c_ulonglong, C.POINTER(cut_config)]),
    "mfg_scene_build": (_vp, [C.POINTER(cut_config)]),
    "mfg_scene_free": (None, [_vp]),
    "mfg_scene_get_params": (_i, [_vp, C.POINTER(scene_params)]),
    "mfg_scene_get_particles": (_i, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mfg_scene_tool_size": (_i, [_vp, _i, _ip, _ip]),
    "mfg_scene_tool_mesh": (_i, [_vp, _i, _dp, _ip]),
    "mfg_mesh_read_ldynak": (_i, [C.c_char_p, _i, _ip, _ip, _dp, _ip]),
    "mfg_mesh_read_json": (_i, [C.c_char_p, _ip, _ip, _dp, _ip]),
    "mfg_mesh_last_error": (C.c_char_p, []),
    "mfg_vtk_read_positions": (_i, [C.c_char_p, _ip, _dp]),
    "mfg_constants_reset": (_i, []),
    "mfg_read_steady_temperatures": (_i, [C.c_char_p, _ip, _dp]),
    "mfg_interp_temps": (_i, [_dp, _i, C.c_double, C.c_double, _dp, _i, _dp]),
}

_lib = None


def load() -> C.CDLL:
    """Load libmfree_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with iwf_mfree_gpu_3d_amd.build.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc < 0:
        msg = load().mfg_last_error()
        raise MfgError(rc, msg.decode() if msg else "")
    return rc


_initialised = False


def init(device: int = 0) -> None:
    """mfg_init on `device` once per process."""
    global _initialised
    if not _initialised:
        check(load().mfg_init(device))
        _initialised = True
