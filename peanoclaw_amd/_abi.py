"""ctypes binding of libpeanoclaw_amd.so (the C ABI in include/peanoclaw_amd.h).

The library is loaded from this package directory.  There is no fallback:
if the HIP library is missing or no gfx950 device is present, every entry
point raises.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# PEANOCLAW_AMD_LIB selects another build of the same library (e.g. a
# variant compiled with different tuning macros)
LIB_PATH = os.environ.get("PEANOCLAW_AMD_LIB") or os.path.join(PKG, "libpeanoclaw_amd.so")

PC_OK = 0
PC_E_INVALID = 1
PC_E_OOM = 2
PC_E_HIP = 3
PC_E_NONFINITE = 4
PC_E_NODEVICE = 5

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class CellDescription(ctypes.Structure):
    """pc_cell_description (mirrors dastgen/CellDescription.def fields)."""
    _fields_ = [
        ("position", ctypes.c_double * 2),
        ("size", ctypes.c_double * 2),
        ("subdivision_factor", ctypes.c_int32 * 2),
        ("ghostlayer_width", ctypes.c_int32),
        ("level", ctypes.c_int32),
        ("is_virtual", ctypes.c_int32),
        ("scenario", ctypes.c_int32),
        ("time", ctypes.c_double),
        ("timestep_size", ctypes.c_double),
        ("estimated_next_timestep_size", ctypes.c_double),
        ("minimal_neighbor_time_constraint", ctypes.c_double),
        ("demanded_mesh_width", ctypes.c_double * 2),
        ("restriction_lower_bounds", ctypes.c_double * 2),
        ("restriction_upper_bounds", ctypes.c_double * 2),
    ]


class GhostStrip(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "dst_slot", "dst_cell", "dst_stride", "src_slot", "src_cell",
        "src_stride", "length", "negate")]


GHOST_STRIP_DTYPE = np.dtype([(n, np.int32) for n in (
    "dst_slot", "dst_cell", "dst_stride", "src_slot", "src_cell",
    "src_stride", "length", "negate")])


MAX_HILLS = 8


class Request(ctypes.Structure):
    """pc_request: one dam-break scenario of the served stream."""
    _fields_ = [
        ("cells_per_side", ctypes.c_int32),
        ("patch_cells", ctypes.c_int32),
        ("end_time", ctypes.c_double),
        ("column_x", ctypes.c_double),
        ("column_y", ctypes.c_double),
        ("column_radius", ctypes.c_double),
        ("column_height", ctypes.c_double),
        ("depth", ctypes.c_double),
        ("num_hills", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("hill_x", ctypes.c_double * MAX_HILLS),
        ("hill_y", ctypes.c_double * MAX_HILLS),
        ("hill_sigma", ctypes.c_double * MAX_HILLS),
        ("hill_height", ctypes.c_double * MAX_HILLS),
    ]


class RequestResult(ctypes.Structure):
    """pc_request_result."""
    _fields_ = [
        ("id", ctypes.c_int32),
        ("scenario", ctypes.c_int32),
        ("steps", ctypes.c_longlong),
        ("cell_updates", ctypes.c_longlong),
        ("time", ctypes.c_double),
        ("mass", ctypes.c_double),
        ("state_hash", ctypes.c_uint64),
        ("admitted_round", ctypes.c_longlong),
        ("completed_round", ctypes.c_longlong),
    ]


class AmrGhost(ctypes.Structure):
    """pc_amr_ghost: kind 0 transfer, 1 interpolation."""
    _fields_ = [("kind", ctypes.c_int32), ("src", ctypes.c_int32), ("dst", ctypes.c_int32),
                ("size", ctypes.c_int32 * 2), ("src_offset", ctypes.c_int32 * 2),
                ("dst_offset", ctypes.c_int32 * 2)]


class AmrWall(ctypes.Structure):
    _fields_ = [("slot", ctypes.c_int32), ("dimension", ctypes.c_int32),
                ("upper", ctypes.c_int32)]


class PeanoClawAmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"peanoclaw_amd error {code}: {msg}")
        self.code = code


_SIGS = {
    "pc_last_error_string": (ctypes.c_char_p, []),
    "pc_pool_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t,
                                      ctypes.POINTER(ctypes.c_void_p)]),
    "pc_pool_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "pc_pool_bytes_in_use": (ctypes.c_int, [ctypes.c_void_p,
                                            ctypes.POINTER(ctypes.c_size_t)]),
    "pc_pool_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "pc_pool_reserve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong]),
    "pc_patch_register": (ctypes.c_int, [ctypes.c_void_p,
                                         ctypes.POINTER(CellDescription), _ip]),
    "pc_patch_release": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "pc_patch_get_description": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                                ctypes.POINTER(CellDescription)]),
    "pc_patch_set_description": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                                ctypes.POINTER(CellDescription)]),
    "pc_patch_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]),
    "pc_patch_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]),
    "pc_patch_offset": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_longlong)]),
    "pc_patch_stride": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip]),
    "pc_tile_stride": (ctypes.c_int, [ctypes.c_int]),
    "pc_scenario_current_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip]),
    "pc_pool_copy_fields": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_longlong, ctypes.c_longlong,
                                           _dp, _dp, _dp, _dp]),
    "pc_solve_timestep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                         _ip, _dp, _dp]),
    "pc_patch_cfl": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "pc_fill_boundary_layer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int]),
    "pc_transfer_ghostlayer": (ctypes.c_int, [ctypes.c_void_p, _ip, _ip, _ip,
                                              ctypes.c_int, ctypes.c_int]),
    "pc_interpolate_solution": (ctypes.c_int, [ctypes.c_void_p, _ip, _ip, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int]),
    "pc_restrict_solution": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int]),
    "pc_post_process_restriction": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                                   ctypes.c_int]),
    "pc_interpolate_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _ip, _ip, _ip,
                                            _ip]),
    "pc_restrict_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _ip, ctypes.c_int]),
    "pc_switch_to_leaf": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _dp]),
    "pc_patches_clear": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip]),
    "pc_patches_set_b": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _dp]),
    "pc_copy_unew_to_uold": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "pc_swap_unew_and_uold": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "pc_demanded_mesh_width": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                              ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, _dp]),
    "pc_get_number_of_unknowns_per_cell": (ctypes.c_int, []),
    "pc_get_number_of_parameter_fields_without_ghostlayer": (ctypes.c_int, []),
    "pc_get_number_of_parameter_fields_with_ghostlayer": (ctypes.c_int, []),
    "pc_get_ghostlayer_width": (ctypes.c_int, []),
    "pc_set_ghost_strips": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "pc_grid_strips": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip,
                                      ctypes.c_void_p, ctypes.c_int, _ip]),
    "pc_set_grid_sides": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip]),
    "pc_clear_sides": (ctypes.c_int, [ctypes.c_void_p]),
    "pc_scenario_set": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double]),
    "pc_scenario_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp,
                                       ctypes.POINTER(ctypes.c_longlong)]),
    "pc_set_launch_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int]),
    "pc_set_prune_slack": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "pc_set_row_segments": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "pc_set_wave_propagation_solver": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "pc_step_all": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "pc_step_all_per_patch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "pc_graph_capture": (ctypes.c_int, [ctypes.c_void_p]),
    "pc_graph_replay": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "pc_refinement_flags": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, _ip, ctypes.c_int]),
    "pc_pool_slot_capacity": (ctypes.c_int, [ctypes.c_void_p, _ip]),
    "pc_scenario_mass": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "pc_last_timing": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "pc_edge_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _dp, _dp]),
    "pc_edge_batch_static": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, _dp, _dp]),
    "pc_compute_fluxes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "pc_apply_flux_correction": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int]),
    "pc_patch_fluxes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]),
    "pc_copy_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, _dp]),
    "pc_vtu_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p]),
    "pc_vtk_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p]),
    "pc_amr_lts_init": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pc_amr_lts_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_longlong)]),
    "pc_pool_write_vtu": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p,
                                         ctypes.c_int]),
    "pc_pool_write_vtk": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p,
                                         ctypes.c_int]),
    "pc_stream_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, _dp]),
    "pc_march_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "pc_server_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "pc_server_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "pc_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Request), _ip]),
    "pc_server_round": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _ip]),
    "pc_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(RequestResult),
                                      ctypes.c_int, _ip]),
    "pc_server_request_slots": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip,
                                               ctypes.c_int, _ip]),
    "pc_server_release": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "pc_pool_device_fields": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] +
                              [ctypes.POINTER(_dp)] * 4),
    "pc_amr_face_plan": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_int, _ip, ctypes.c_void_p, ctypes.c_int, _ip]),
    "pc_amr_create": (ctypes.c_int, [ctypes.c_void_p, _ip, ctypes.c_int, _ip, ctypes.c_int,
                                     _ip, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_void_p)]),
    "pc_amr_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp]),
    "pc_amr_clock": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp,
                                    ctypes.POINTER(ctypes.c_longlong)]),
    "pc_amr_destroy": (ctypes.c_int, [ctypes.c_void_p]),
}

EXPORTED_SYMBOLS = tuple(_SIGS)

_lib = None


def load():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PeanoClawAmdError(
                PC_E_NODEVICE,
                f"{LIB_PATH} is missing; build it with python -m peanoclaw_amd.build")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    if rc != PC_OK:
        msg = load().pc_last_error_string().decode(errors="replace")
        raise PeanoClawAmdError(rc, msg)


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iarr(*v):
    return (ctypes.c_int * len(v))(*v)
