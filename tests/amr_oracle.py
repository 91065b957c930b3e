"""Oracle backend of the two-level AMR driver (test infrastructure): the same
Numerics calls as peanoclaw_amd.amr.DeviceBackend, run by the C oracle on
patches in the reference's layout."""
import ctypes

import numpy as np

from oracle_lib import Patch, iarr, lib


class OracleBackend:
    def __init__(self):
        self.p = []

    def new_patch(self, pos, size, n, level, virtual, time=0.0, dt=0.0):
        P = Patch(3, 1, n, 1, pos, size, level, time, dt)
        P.s.is_virtual = 1 if virtual else 0
        P.s.is_leaf = 0 if virtual else 1
        self.p.append(P)
        return len(self.p) - 1

    def set_state(self, s, h, b):
        P = self.p[s]
        P.unew[:] = 0.0
        P.unew[0] = h
        P.uold[:] = 0.0
        P.pwg[0] = b

    def set_b(self, s, b):
        self.p[s].pwg[0] = b

    def set_virtual(self, s, virtual):
        self.p[s].s.is_virtual = 1 if virtual else 0
        self.p[s].s.is_leaf = 0 if virtual else 1

    def release(self, s):
        self.p[s] = None

    def flags(self, slots, threshold, minimal_width, maximal_width):
        out = []
        for s in slots:
            P = self.p[s]
            dw = (ctypes.c_double * 2)()
            lib().or_demanded_mesh_width(P.ref, threshold, minimal_width, maximal_width, dw)
            sub = (ctypes.c_double * 2)(*P.s.subcell)
            out.append(lib().or_refinement_decision(sub, dw))
        return out

    def switch_to_leaf(self, s, time, dt):
        P = self.p[s]
        P.s.time, P.s.dt = time, dt
        lib().or_switch_to_leaf(P.ref)

    def clear(self, s):
        self.p[s].unew[:] = 0.0
        self.p[s].uold[:] = 0.0

    def restrict(self, src, dst):
        lib().or_restrict_swe(self.p[src].ref, self.p[dst].ref, 0)

    def restrict_pure(self, src, dst):
        lib().or_restrict_swe_pure(self.p[src].ref, self.p[dst].ref)

    def interpolate_pure(self, size, doff, src, dst, to_uold, to_current, use_time):
        if lib().or_interpolate_swe_pure(iarr(*size), iarr(*doff), self.p[src].ref,
                                         self.p[dst].ref, to_uold, to_current, use_time):
            raise RuntimeError("interpolation region outside its patch")

    def post_restrict(self, dst):
        lib().or_post_process_restriction_swe(self.p[dst].ref, 0)

    def copy_unew_to_uold(self, s):
        lib().or_copy_unew_to_uold(self.p[s].ref)

    def transfer(self, size, soff, doff, src, dst):
        lib().or_transfer_ghostlayer(iarr(*size), iarr(*soff), iarr(*doff), self.p[src].ref,
                                     self.p[dst].ref)

    def interpolate(self, size, doff, src, dst, to_uold, to_current, use_time):
        if lib().or_interpolate_swe(iarr(*size), iarr(*doff), self.p[src].ref, self.p[dst].ref,
                                    to_uold, to_current, use_time):
            raise RuntimeError("interpolation region outside its patch")

    # batched forms: the per-call methods one after another
    def interpolate_batch(self, calls):
        for c in calls:
            self.interpolate(*c)

    def restrict_batch(self, pairs):
        for src, dst in pairs:
            self.restrict(src, dst)

    def clear_batch(self, handles):
        for s in handles:
            self.clear(s)

    def set_b_batch(self, handles, bs):
        for s, b in zip(handles, bs):
            self.set_b(s, b)

    def switch_to_leaf_batch(self, handles, time, dt):
        for s in handles:
            self.switch_to_leaf(s, time, dt)

    def fill_boundary(self, s, dim, upper):
        lib().or_fill_boundary_layer(self.p[s].ref, dim, upper)

    def cfl(self, s):
        return lib().or_patch_cfl(self.p[s].ref)

    def solve(self, s, max_dt):
        return lib().or_solve_timestep(self.p[s].ref, float(max_dt))

    def clock(self, s):
        return self.p[s].s.time, self.p[s].s.dt

    def state(self, s):
        return np.array(self.p[s].unew)

    def uold(self, s):
        return np.array(self.p[s].uold)
