"""The Numerics plugin interface of PeanoClaw for the SWE kernel, on the GPU.

Same method names, argument meaning and side effects as
peanoclaw::Numerics (src/peanoclaw/Numerics.h:30-204) implemented by
peanoclaw::native::SWEKernel (src/peanoclaw/native/SWEKernel.cpp), with a
`Patch` handle (pool + slot) in place of the heap-backed peanoclaw::Patch.
Every call goes through the C ABI to the HIP kernels; there is no host
fallback.
"""
import numpy as np

from ._abi import check, dptr, iarr
from .pool import PatchPool, make_description


class Patch:
    """Facade over one registered slot (peanoclaw::Patch, Patch.h:135-516)."""

    def __init__(self, pool, slot):
        self.pool = pool
        self.slot = slot

    @property
    def description(self):
        return self.pool.description(self.slot)

    def getSubdivisionFactor(self):
        return self.pool.dims(self.slot)

    def getSubcellSize(self):
        d = self.description
        return (d.size[0] / d.subdivision_factor[0], d.size[1] / d.subdivision_factor[1])

    def getPosition(self):
        d = self.description
        return (d.position[0], d.position[1])

    def getTimeIntervals(self):
        d = self.description
        return {"time": d.time, "timestepSize": d.timestep_size,
                "estimatedNextTimestepSize": d.estimated_next_timestep_size}

    def arrays(self):
        """(uNew, uOldWithGhostlayer, parameterWithGhostlayer) as numpy."""
        return self.pool.download(self.slot)


class BreakingDamSWEScenario:
    """native/scenarios/BreakingDam.cpp:53-139 with the dam parameters of
    the configuration (centre, radius, inner/outer depth) and the gradient
    threshold of computeDemandedMeshWidth."""

    def __init__(self, minimal_mesh_width, maximal_mesh_width, centre=(0.5, 0.5),
                 radius=0.25, h_in=2.0, h_out=1.0, threshold=0.1):
        self.minimal_mesh_width = float(minimal_mesh_width)
        self.maximal_mesh_width = float(maximal_mesh_width)
        self.centre = centre
        self.radius = radius
        self.h_in = h_in
        self.h_out = h_out
        self.threshold = threshold

    def water_height(self, X, Y):
        r = np.sqrt((X - self.centre[0]) * (X - self.centre[0]) +
                    (Y - self.centre[1]) * (Y - self.centre[1]))
        return self.h_in * (r <= self.radius) + self.h_out * (r > self.radius)

    def initializePatch(self, patch):
        d = patch.description
        nx, ny = d.subdivision_factor[0], d.subdivision_factor[1]
        dx, dy = d.size[0] / nx, d.size[1] / ny
        X = d.position[0] + (np.arange(nx)[:, None] + 0.5) * dx
        Y = d.position[1] + (np.arange(ny)[None, :] + 0.5) * dy
        unew = np.zeros((3, nx, ny))
        unew[0] = self.water_height(X, Y)
        b = np.zeros((nx + 2, ny + 2))
        patch.pool.upload(patch.slot, unew=unew, b=b)


class SWENumerics:
    """peanoclaw::native::SWEKernel on the GPU (SWEKernel.h:32-170)."""

    def __init__(self, pool: PatchPool, scenario=None):
        self.pool = pool
        self.scenario = scenario
        self.lib = pool.lib

    # Numerics.h:58 / SWEKernel.cpp:111-119
    def initializePatch(self, patch):
        self.scenario.initializePatch(patch)

    # Numerics.h:169 / SWEKernel.cpp:121-123 (BreakingDam: b stays 0)
    def update(self, patch):
        pass

    # Numerics.h:147-152 / SWEKernel.cpp:125-167
    def solveTimestep(self, patch, maximumTimestepSize, useDimensionalSplitting=False,
                      domainBoundaryFlags=(False, False, False, False)):
        import ctypes
        if not maximumTimestepSize > 0.0:
            raise ValueError("Timestepsize == 0 should be checked outside.")
        dt, est = ctypes.c_double(), ctypes.c_double()
        flags = iarr(*[1 if f else 0 for f in domainBoundaryFlags])
        check(self.lib.pc_solve_timestep(self.pool.h, patch.slot, float(maximumTimestepSize),
                                         flags, ctypes.byref(dt), ctypes.byref(est)))
        return dt.value, est.value

    # Numerics.h:131-135 / SWEKernel.cpp:176-237
    def fillBoundaryLayer(self, patch, dimension, setUpper):
        check(self.lib.pc_fill_boundary_layer(self.pool.h, patch.slot, int(dimension),
                                              1 if setUpper else 0))

    # Numerics.h:64-70 / DefaultTransfer.cpph:15-101
    def transferGhostlayer(self, size, sourceOffset, destinationOffset, source, destination):
        check(self.lib.pc_transfer_ghostlayer(self.pool.h, iarr(*size), iarr(*sourceOffset),
                                              iarr(*destinationOffset), source.slot,
                                              destination.slot))

    # Numerics.h:77-85 / SWEKernel.cpp:239-286
    def interpolateSolution(self, destinationSize, destinationOffset, source, destination,
                            interpolateToUOld, interpolateToCurrentTime,
                            useTimeUNewOrTimeUOld):
        check(self.lib.pc_interpolate_solution(
            self.pool.h, iarr(*destinationSize), iarr(*destinationOffset), source.slot,
            destination.slot, int(bool(interpolateToUOld)), int(bool(interpolateToCurrentTime)),
            int(bool(useTimeUNewOrTimeUOld))))

    # Numerics.h:92-96 / SWEKernel.cpp:288-307
    def restrictSolution(self, source, destination, restrictOnlyOverlappedRegions):
        check(self.lib.pc_restrict_solution(self.pool.h, source.slot, destination.slot,
                                            int(bool(restrictOnlyOverlappedRegions))))

    # Numerics.h:103-106 / SWEKernel.cpp:309-316
    def postProcessRestriction(self, destination, restrictOnlyOverlappedRegions):
        check(self.lib.pc_post_process_restriction(self.pool.h, destination.slot,
                                                   int(bool(restrictOnlyOverlappedRegions))))

    # Numerics.h:160-163 / SWEKernel.cpp:169-171 -> BreakingDam.cpp:85-139
    def getDemandedMeshWidth(self, patch, isInitializing=False):
        import ctypes
        out = (ctypes.c_double * 2)()
        s = self.scenario
        check(self.lib.pc_demanded_mesh_width(self.pool.h, patch.slot, s.threshold,
                                              s.minimal_mesh_width, s.maximal_mesh_width, out))
        return (out[0], out[1])

    # Numerics.cpp:93-98 / DefaultFluxCorrection.cpph:44-86 (SolveTimestep.cpp:673-679)
    def computeFluxes(self, subgrid):
        check(self.lib.pc_compute_fluxes(self.pool.h, subgrid.slot))

    # Numerics.cpp:80-91 / DefaultFluxCorrection.cpph:88-164
    def applyFluxCorrection(self, sourceSubgrid, destinationSubgrid, dimension, direction):
        check(self.lib.pc_apply_flux_correction(self.pool.h, sourceSubgrid.slot,
                                                destinationSubgrid.slot, int(dimension),
                                                int(direction)))

    def fluxes(self, patch, values=None):
        """The patch's flux array (6 (nx + ny) doubles, faces x-lower, x-upper,
        y-lower, y-upper, unknown slowest); uploads `values` when given."""
        nx, ny = self.pool.dims(patch.slot)
        if values is not None:
            v = np.ascontiguousarray(values, dtype=np.float64)
            assert v.size == 6 * (nx + ny)
            check(self.lib.pc_patch_fluxes(self.pool.h, patch.slot, 1, dptr(v)))
            return v
        out = np.empty(6 * (nx + ny))
        check(self.lib.pc_patch_fluxes(self.pool.h, patch.slot, 0, dptr(out)))
        return out

    # DefaultTransfer::copyUNewToUOld (called by SolveTimestep.cpp:638)
    def copyUNewToUOld(self, patch):
        check(self.lib.pc_copy_unew_to_uold(self.pool.h, patch.slot))

    def swapUNewAndUOld(self, patch):
        check(self.lib.pc_swap_unew_and_uold(self.pool.h, patch.slot))

    # SWEKernel.h:124-133
    def getNumberOfUnknownsPerCell(self):
        return self.lib.pc_get_number_of_unknowns_per_cell()

    def getNumberOfParameterFieldsWithoutGhostlayer(self):
        return self.lib.pc_get_number_of_parameter_fields_without_ghostlayer()

    def getNumberOfParameterFieldsWithGhostlayer(self):
        return self.lib.pc_get_number_of_parameter_fields_with_ghostlayer()

    def getGhostlayerWidth(self):
        return self.lib.pc_get_ghostlayer_width()


def create_swe_numerics(pool, scenario):
    """NumericsFactory::createSWENumerics (NumericsFactory.cpp:27-51)."""
    return SWENumerics(pool, scenario)


__all__ = ["Patch", "SWENumerics", "BreakingDamSWEScenario", "create_swe_numerics",
           "make_description"]
