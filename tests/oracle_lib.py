"""ctypes view of the C oracle (test infrastructure; checker only).

The oracle is loaded from ``oracle/liboracle.so``.  Only tests, the smoke()
entry point and bench.py's CPU-baseline leg use it.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrPatch(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("gw", ctypes.c_int),
        ("nq", ctypes.c_int), ("npw", ctypes.c_int),
        ("pos", ctypes.c_double * 2), ("size", ctypes.c_double * 2),
        ("subcell", ctypes.c_double * 2),
        ("level", ctypes.c_int),
        ("is_virtual", ctypes.c_int), ("is_leaf", ctypes.c_int),
        ("is_valid", ctypes.c_int),
        ("time", ctypes.c_double), ("dt", ctypes.c_double),
        ("est_next", ctypes.c_double),
        ("min_neighbor_time_constraint", ctypes.c_double),
        ("min_leaf_neighbor_time_constraint", ctypes.c_double),
        ("min_neighbor_time", ctypes.c_double),
        ("max_neighbor_dt", ctypes.c_double),
        ("lower_ghost_bounds", ctypes.c_double * 2),
        ("upper_ghost_bounds", ctypes.c_double * 2),
        ("cell_description_index", ctypes.c_int),
        ("unew", _dp), ("uold", _dp), ("pwg", _dp),
    ]


class OrTraversalCall(ctypes.Structure):
    _fields_ = [("src", ctypes.c_int), ("dst", ctypes.c_int),
                ("dir", ctypes.c_int * 2)]


class OrGhostEntry(ctypes.Structure):
    _fields_ = [("dst", ctypes.c_int), ("dst_cell", ctypes.c_int),
                ("src", ctypes.c_int), ("src_cell", ctypes.c_int),
                ("negate", ctypes.c_int)]


class OrPool(ctypes.Structure):
    _fields_ = [
        ("np", ctypes.c_int),
        ("nx", _ip), ("ny", _ip),
        ("off", ctypes.POINTER(ctypes.c_longlong)),
        ("dx", _dp), ("dy", _dp),
        ("h", _dp * 2), ("hu", _dp * 2), ("hv", _dp * 2),
        ("b", _dp), ("cur", ctypes.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        P = ctypes.POINTER(OrPatch)
        sig = {
            "or_fwave": (None, [ctypes.c_double] * 6 + [_dp]),
            "or_fwave_speed": (ctypes.c_double, [ctypes.c_double] * 4),
            "or_fwave_batch": (None, [ctypes.c_longlong, _dp, _dp]),
            "or_aug_middle_state": (None, [ctypes.c_double] * 4 + [ctypes.c_int, _dp]),
            "or_aug_waves": (ctypes.c_int, [ctypes.c_double] * 6 + [_dp] * 3),
            "or_augrie": (None, [ctypes.c_double] * 6 + [_dp]),
            "or_set_wave_propagation_solver": (ctypes.c_int, [ctypes.c_int]),
            "or_edge": (None, [ctypes.c_double] * 6 + [_dp]),
            "or_edge_speed": (ctypes.c_double, [ctypes.c_double] * 6),
            "or_edge_batch": (None, [ctypes.c_longlong, _dp, _dp]),
            "or_block_fluxes": (ctypes.c_double, [ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.c_double] + [_dp] * 7),
            "or_block_max_timestep": (ctypes.c_double, [ctypes.c_double] * 3),
            "or_block_update": (None, [ctypes.c_int, ctypes.c_int,
                                ctypes.c_double] + [_dp] * 6),
            "or_patch_init": (None, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, ctypes.c_double,
                              ctypes.c_double, ctypes.c_double, ctypes.c_double,
                              ctypes.c_int, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, _dp, _dp, _dp]),
            "or_solve_timestep": (ctypes.c_double, [P, ctypes.c_double]),
            "or_patch_cfl": (ctypes.c_double, [P]),
            "or_fill_boundary_layer": (None, [P, ctypes.c_int, ctypes.c_int]),
            "or_copy_unew_to_uold": (None, [P]),
            "or_swap_unew_uold": (None, [P]),
            "or_transfer_ghostlayer": (None, [_ip, _ip, _ip, P, P]),
            "or_interpolate_default": (None, [_ip, _ip, P, P, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int]),
            "or_interpolate_swe": (ctypes.c_int, [_ip, _ip, P, P, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int]),
            "or_restrict_default": (None, [P, P, ctypes.c_int]),
            "or_restrict_swe": (None, [P, P, ctypes.c_int]),
            "or_restrict_swe_pure": (None, [P, P]),
            "or_interpolate_swe_pure": (ctypes.c_int, [_ip, _ip, P, P, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int]),
            "or_post_process_restriction_swe": (None, [P, ctypes.c_int]),
            "or_switch_to_leaf": (None, [P]),
            "or_flux_face_offset": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
            "or_flux_array_size": (ctypes.c_int, [P]),
            "or_compute_fluxes": (None, [P, _dp]),
            "or_apply_flux_correction": (ctypes.c_int, [P, _dp, P, _dp, ctypes.c_int,
                                                        ctypes.c_int]),
            "or_overlapping_region": (ctypes.c_double, [_dp] * 4),
            "or_region_map_to_patch": (None, [_ip, _ip, P, P, ctypes.c_double,
                                       _ip, _ip]),
            "or_regions_overlapped_by_ghostlayers": (ctypes.c_int, [_dp] * 5 +
                                                     [_ip, _ip, _ip]),
            "or_face_traversal_order": (ctypes.c_int,
                                        [ctypes.POINTER(OrTraversalCall)]),
            "or_edge_traversal_order": (ctypes.c_int,
                                        [ctypes.POINTER(OrTraversalCall)]),
            "or_should_transfer": (ctypes.c_int, [P, P]),
            "or_fill_ghostlayers_and_update_neighbor_times": (
                None, [ctypes.POINTER(P), ctypes.c_int, ctypes.c_int,
                       ctypes.c_int]),
            "or_is_blocked_by_neighbors": (ctypes.c_int, [P]),
            "or_reset_min_neighbor_time_constraint": (None, [P]),
            "or_minimal_fine_grid_interval": (None, [_dp, _dp, ctypes.c_int,
                                              _dp]),
            "or_breaking_dam_init": (None, [P] + [ctypes.c_double] * 5),
            "or_demanded_mesh_width": (None, [P] + [ctypes.c_double] * 3 + [_dp]),
            "or_refinement_decision": (ctypes.c_int, [_dp, _dp]),
            "or_grid_fill": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip,
                             ctypes.POINTER(ctypes.c_longlong), _dp, _dp, _dp]),
            "or_grid_fill_rect": (None, [ctypes.c_int] * 4 + [_ip,
                                  ctypes.POINTER(ctypes.c_longlong), _dp, _dp, _dp]),
            "or_pool_fill_ghosts": (None, [ctypes.POINTER(OrPool),
                                    ctypes.POINTER(OrGhostEntry),
                                    ctypes.c_longlong]),
            "or_pool_step": (ctypes.c_double, [ctypes.POINTER(OrPool),
                             ctypes.c_double, ctypes.c_int]),
            "or_pool_max_speeds": (None, [ctypes.POINTER(OrPool), ctypes.c_int,
                                   _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def augrie(hL, hR, huL, huR, bL, bR):
    """(dhL, dhR, dqL, dqR, max speed) of the oracle's augmented solver."""
    o = (ctypes.c_double * 5)()
    lib().or_augrie(hL, hR, huL, huR, bL, bR, o)
    return tuple(o)


def aug_waves(hL, hR, huL, huR, bL, bR):
    sw, fh, fq = (ctypes.c_double * 3)(), (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
    lib().or_aug_waves(hL, hR, huL, huR, bL, bR, sw, fh, fq)
    return tuple(sw), tuple(fh), tuple(fq)


def aug_middle_state(hL, hR, uL, uR, maxiter=1):
    o = (ctypes.c_double * 5)()
    lib().or_aug_middle_state(hL, hR, uL, uR, maxiter, o)
    return tuple(o)


class wave_propagation_solver:
    """Context manager selecting the oracle's edge solver (1, 4 f-wave;
    2 augmented Riemann), restoring the f-wave on exit."""

    def __init__(self, solver):
        self.solver = solver

    def __enter__(self):
        assert lib().or_set_wave_propagation_solver(self.solver) == 0
        return self

    def __exit__(self, *exc):
        lib().or_set_wave_propagation_solver(1)


def clamp_events():
    """Cells whose depth the oracle's update has clamped from below zero."""
    return ctypes.c_longlong.in_dll(lib(), "or_clamp_events").value


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iarr(*v):
    return (ctypes.c_int * len(v))(*v)


def darr(*v):
    return (ctypes.c_double * len(v))(*v)


class Patch:
    """An oracle patch with numpy-owned arrays in the reference layout."""

    def __init__(self, nq, npw, n, gw, pos, size, level, time, dt,
                 min_neighbor_time=0.0, ny=None):
        nx = n
        ny = n if ny is None else ny
        self.unew = np.zeros((nq, nx, ny))
        self.uold = np.zeros((nq, nx + 2 * gw, ny + 2 * gw))
        self.pwg = np.zeros((max(npw, 1), nx + 2 * gw, ny + 2 * gw))
        self.s = OrPatch()
        px, py = (pos, pos) if np.isscalar(pos) else pos
        sx, sy = (size, size) if np.isscalar(size) else size
        lib().or_patch_init(ctypes.byref(self.s), nq, npw, nx, ny, gw,
                            float(px), float(py), float(sx), float(sy), level,
                            float(time), float(dt), float(min_neighbor_time),
                            dptr(self.unew), dptr(self.uold), dptr(self.pwg))
        self.gw = gw

    @property
    def ref(self):
        return ctypes.byref(self.s)

    def uold_at(self, q, x, y):
        return self.uold[q, x + self.gw, y + self.gw]

    def set_uold(self, q, x, y, v):
        self.uold[q, x + self.gw, y + self.gw] = v


def invalid_patch():
    p = Patch(1, 0, 2, 1, 0.0, 1.0, 0, 0.0, 0.0)
    p.s.is_valid = 0
    p.s.is_leaf = 0
    return p


def fwave(hL, hR, huL, huR, bL, bR):
    out = (ctypes.c_double * 5)()
    lib().or_fwave(hL, hR, huL, huR, bL, bR, out)
    return tuple(out)
