"""Host-side patch pool over the C ABI (device memory lives in the library).

`PatchPool` replaces Peano's DataHeap/CellDescriptionHeap for the SWE path:
patches are registered with a CellDescription-shaped record and addressed
by slot id.  Arrays exchanged with a slot use the reference's layout
(grid/LinearizationZYXQ.h:24-82): q slowest, x, then y fastest.
"""
import ctypes

import numpy as np

from ._abi import (CellDescription, GHOST_STRIP_DTYPE, check, dptr, iarr, load)


def make_description(position, size, n, scenario=0, level=0, time=0.0,
                     timestep_size=0.0, is_virtual=False, ny=None):
    d = CellDescription()
    px, py = (position, position) if np.isscalar(position) else position
    sx, sy = (size, size) if np.isscalar(size) else size
    d.position[0], d.position[1] = float(px), float(py)
    d.size[0], d.size[1] = float(sx), float(sy)
    d.subdivision_factor[0] = int(n)
    d.subdivision_factor[1] = int(n if ny is None else ny)
    d.ghostlayer_width = 1
    d.level = level
    d.is_virtual = 1 if is_virtual else 0
    d.scenario = scenario
    d.time = time
    d.timestep_size = timestep_size
    d.minimal_neighbor_time_constraint = np.finfo(np.float64).max
    d.restriction_lower_bounds[0] = d.restriction_lower_bounds[1] = np.finfo(np.float64).max
    d.restriction_upper_bounds[0] = d.restriction_upper_bounds[1] = -np.finfo(np.float64).max
    return d


def tile_stride(ny):
    """Row stride (doubles) of the pool's tiles with ny interior columns:
    ny + 2, padded to a multiple of 16 for ny >= 8 so that every row's
    interior starts on a 128-byte line."""
    return load().pc_tile_stride(int(ny))


class PatchPool:
    """A pool of SWE patches resident in HBM on one gfx950 device."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.pc_pool_create(device, hbm_budget_bytes, ctypes.byref(h)))
        self.h = h
        self._dims = {}

    def close(self):
        if self.h:
            check(self.lib.pc_pool_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- registry ----
    def register(self, desc):
        s = ctypes.c_int()
        check(self.lib.pc_patch_register(self.h, ctypes.byref(desc), ctypes.byref(s)))
        self._dims[s.value] = (desc.subdivision_factor[0], desc.subdivision_factor[1])
        return s.value

    def release(self, slot):
        check(self.lib.pc_patch_release(self.h, slot))
        self._dims.pop(slot, None)

    def description(self, slot):
        d = CellDescription()
        check(self.lib.pc_patch_get_description(self.h, slot, ctypes.byref(d)))
        return d

    def set_description(self, slot, d):
        check(self.lib.pc_patch_set_description(self.h, slot, ctypes.byref(d)))

    def dims(self, slot):
        return self._dims[slot]

    def offset(self, slot):
        o = ctypes.c_longlong()
        check(self.lib.pc_patch_offset(self.h, slot, ctypes.byref(o)))
        return o.value

    def stride(self, slot):
        """Row stride of the slot's tile in the field arrays (row x of the
        tile, ghost rows included, starts at offset(slot) + x * stride)."""
        o = ctypes.c_int()
        check(self.lib.pc_patch_stride(self.h, slot, ctypes.byref(o)))
        return o.value

    # ---- data ----
    def upload(self, slot, unew=None, uold=None, b=None):
        def c(a):
            return None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        unew, uold, b = c(unew), c(uold), c(b)
        check(self.lib.pc_patch_upload(self.h, slot, dptr(unew), dptr(uold), dptr(b)))

    def download(self, slot):
        if slot not in self._dims:  # registered inside the library (e.g. by a server)
            d = self.description(slot)
            self._dims[slot] = (d.subdivision_factor[0], d.subdivision_factor[1])
        nx, ny = self._dims[slot]
        unew = np.empty((3, nx, ny))
        uold = np.empty((3, nx + 2, ny + 2))
        b = np.empty((nx + 2, ny + 2))
        check(self.lib.pc_patch_download(self.h, slot, dptr(unew), dptr(uold), dptr(b)))
        return unew, uold, b

    def current_state(self, scenario=0):
        s = ctypes.c_int()
        check(self.lib.pc_scenario_current_state(self.h, scenario, ctypes.byref(s)))
        return s.value

    def copy_fields(self, to_device, state, begin, h, hu, hv, b=None):
        n = h.size
        check(self.lib.pc_pool_copy_fields(self.h, 1 if to_device else 0, state, begin, n,
                                           dptr(h), dptr(hu), dptr(hv), dptr(b)))

    # ---- ghost table ----
    def grid_strips(self, slot_of):
        slot_of = np.ascontiguousarray(slot_of, dtype=np.int32)
        px, py = slot_of.shape
        n = ctypes.c_int()
        ip = slot_of.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        check(self.lib.pc_grid_strips(self.h, px, py, ip, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=GHOST_STRIP_DTYPE)
        check(self.lib.pc_grid_strips(self.h, px, py, ip, out.ctypes.data_as(ctypes.c_void_p),
                                      n.value, ctypes.byref(n)))
        return out

    def set_grid_sides(self, slot_of):
        slot_of = np.ascontiguousarray(slot_of, dtype=np.int32)
        px, py = slot_of.shape
        check(self.lib.pc_set_grid_sides(self.h, px, py,
                                         slot_of.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def clear_sides(self):
        check(self.lib.pc_clear_sides(self.h))

    def set_ghost_strips(self, strips):
        strips = np.ascontiguousarray(strips, dtype=GHOST_STRIP_DTYPE)
        check(self.lib.pc_set_ghost_strips(self.h, strips.ctypes.data_as(ctypes.c_void_p),
                                           len(strips)))

    # ---- scenarios and stepping ----
    def scenario_set(self, scenario, time=0.0, timestep_size=0.0, end_time=np.inf):
        check(self.lib.pc_scenario_set(self.h, scenario, time, timestep_size, end_time))

    def scenario_get(self, scenario=0):
        t, dt, est = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_longlong()
        check(self.lib.pc_scenario_get(self.h, scenario, ctypes.byref(t), ctypes.byref(dt),
                                       ctypes.byref(est), ctypes.byref(n)))
        return t.value, dt.value, est.value, n.value

    def launch_config(self, persistent_workgroups=0, fused=0, patches_per_task_cap=0):
        check(self.lib.pc_set_launch_config(self.h, persistent_workgroups, fused,
                                            patches_per_task_cap))

    def wave_propagation_solver(self, solver):
        """Edge solver of the pool's updates: 1 or 4 f-wave, 2 augmented
        Riemann (the reference's solverType build option)."""
        check(self.lib.pc_set_wave_propagation_solver(self.h, int(solver)))

    def prune_slack(self, slack):
        """Threshold slack of the pruned speed mode (fused=2); see the C header."""
        check(self.lib.pc_set_prune_slack(self.h, float(slack)))

    def row_segments(self, segments):
        """Row segments of the batched update: 0 automatic, 1 whole patches,
        2 or 4 forced (see the C header)."""
        check(self.lib.pc_set_row_segments(self.h, int(segments)))

    def step(self, k=1):
        check(self.lib.pc_step_all(self.h, k))

    def step_per_patch(self, k=1):
        check(self.lib.pc_step_all_per_patch(self.h, k))

    def graph_capture(self):
        check(self.lib.pc_graph_capture(self.h))

    def graph_replay(self, k=1):
        check(self.lib.pc_graph_replay(self.h, k))

    def last_timing(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(self.lib.pc_last_timing(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def mass(self, scenario=0):
        m = ctypes.c_double()
        check(self.lib.pc_scenario_mass(self.h, scenario, ctypes.byref(m)))
        return m.value

    def slot_capacity(self):
        n = ctypes.c_int()
        check(self.lib.pc_pool_slot_capacity(self.h, ctypes.byref(n)))
        return n.value

    def refinement_flags(self, threshold, minimal_width, maximal_width, nslots=0):
        """Flags per slot id (-1 unused), at least `nslots` entries."""
        cap = max(int(nslots), self.slot_capacity())
        f = np.empty(cap, dtype=np.int32)
        check(self.lib.pc_refinement_flags(self.h, threshold, minimal_width, maximal_width,
                                           f.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), cap))
        return f

    def reserve(self, cells):
        check(self.lib.pc_pool_reserve(self.h, int(cells)))

    def synchronize(self):
        check(self.lib.pc_pool_synchronize(self.h))

    def bytes_in_use(self):
        n = ctypes.c_size_t()
        check(self.lib.pc_pool_bytes_in_use(self.h, ctypes.byref(n)))
        return n.value

    def edge_batch(self, edges):
        edges = np.ascontiguousarray(edges, dtype=np.float64)
        n = edges.shape[0]
        out = np.empty((n, 5))
        check(self.lib.pc_edge_batch(self.h, n, dptr(edges), dptr(out)))
        return out

    def copy_probe(self, nbytes, reps=5):
        """Best device-to-device copy rate in GB/s (read + written bytes)."""
        v = ctypes.c_double()
        check(self.lib.pc_copy_probe(self.h, int(nbytes), int(reps), ctypes.byref(v)))
        return v.value

    def stream_probe(self, cells, reps=5):
        """Best rate in GB/s of a flat stream with the update's traffic mix
        (4 doubles read, 3 written per element; 56 B x `cells`)."""
        v = ctypes.c_double()
        check(self.lib.pc_stream_probe(self.h, int(cells), int(reps), ctypes.byref(v)))
        return v.value

    def march_probe(self, reps=5):
        """Best rate in GB/s (56 B per interior cell) of the batched update's
        traversal of the pool without the edge arithmetic."""
        v = ctypes.c_double()
        check(self.lib.pc_march_probe(self.h, int(reps), ctypes.byref(v)))
        return v.value


__all__ = ["PatchPool", "make_description", "iarr"]
