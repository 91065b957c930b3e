"""Flux correction on the GPU through the C ABI (pc_compute_fluxes,
pc_apply_flux_correction, pc_patch_fluxes) against the oracle's
restatement of the present reference code (DefaultFluxCorrection.cpph:
13-164), bitwise: fluxes of random patches, then corrections between a
coarse patch and a fine neighbour (factor 3) on every side, in both roles,
with fluxes computed on the device and with uploaded random fluxes."""
import numpy as np
import pytest

from oracle_lib import Patch as OPatch, dptr as odptr, lib as olib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import peanoclaw_amd  # noqa: F401
    from peanoclaw_amd.pool import PatchPool
    pool = PatchPool(0)
    yield pool
    pool.close()


def _pair(amd, rng, n, pos, size, level, time, dt):
    from peanoclaw_amd.pool import make_description
    s = amd.register(make_description(pos, size, n, level=level, time=time, timestep_size=dt))
    unew = np.stack([rng.uniform(0.5, 2.0, (n, n)), rng.uniform(-0.4, 0.4, (n, n)),
                     rng.uniform(-0.4, 0.4, (n, n))])
    uold = np.stack([rng.uniform(0.5, 2.0, (n + 2, n + 2)),
                     rng.uniform(-0.4, 0.4, (n + 2, n + 2)),
                     rng.uniform(-0.4, 0.4, (n + 2, n + 2))])
    b = np.zeros((n + 2, n + 2))
    amd.upload(s, unew=unew, uold=uold, b=b)
    P = OPatch(3, 1, n, 1, pos, size, level, time, dt)
    P.unew[:] = unew
    P.uold[:] = uold
    return s, P


def test_compute_fluxes_bitwise(amd):
    from peanoclaw_amd.numerics import Patch, SWENumerics
    rng = np.random.default_rng(3)
    num = SWENumerics(amd)
    for n in (2, 3, 16, 64):
        s, P = _pair(amd, rng, n, (0.25, 0.5), 0.3, 0, 0.1, 0.01)
        num.computeFluxes(Patch(amd, s))
        got = num.fluxes(Patch(amd, s))
        ref = np.zeros(olib().or_flux_array_size(P.ref))
        olib().or_compute_fluxes(P.ref, odptr(ref))
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), n
        amd.release(s)


@pytest.mark.parametrize("n", [2, 3, 9])
def test_apply_flux_correction_bitwise(amd, n):
    from peanoclaw_amd.numerics import Patch, SWENumerics
    num = SWENumerics(amd)
    fs = 1.0 / 3.0
    cases = 0
    for dim in (0, 1):
        for d in (-1, 1):
            for k in range(3):
                rng = np.random.default_rng(1000 * n + 100 * dim + 10 * (d + 1) + k)
                fpos = [1.0 + k * fs, 1.0 + k * fs]
                fpos[dim] = 1.0 - fs if d == -1 else 2.0
                sc, C = _pair(amd, rng, n, (1.0, 1.0), 1.0, 0, 0.2, 0.05)
                sf, F = _pair(amd, rng, n, tuple(fpos), fs, 1, 0.21, 0.02)
                # fluxes: computed on the device for the coarse patch, uploaded
                # random values for the fine one
                num.computeFluxes(Patch(amd, sc))
                fc = np.zeros(12 * n)
                olib().or_compute_fluxes(C.ref, odptr(fc))
                ff = rng.uniform(-1, 1, 12 * n)
                num.fluxes(Patch(amd, sf), ff)
                for (ss, S, sflux), (ds, D, dflux), dd in (
                        ((sc, C, fc), (sf, F, ff), d), ((sf, F, ff), (sc, C, fc), -d)):
                    num.applyFluxCorrection(Patch(amd, ss), Patch(amd, ds), dim, dd)
                    assert olib().or_apply_flux_correction(S.ref, odptr(sflux), D.ref,
                                                           odptr(dflux), dim, dd) == 0
                    got = amd.download(ds)[0]
                    assert np.array_equal(got.view(np.uint64),
                                          np.array(D.unew).view(np.uint64)), (dim, d, k)
                    cases += 1
                amd.release(sc)
                amd.release(sf)
    assert cases == 24


def test_flux_correction_rejects_what_the_reference_leaves_undefined(amd):
    from peanoclaw_amd._abi import PeanoClawAmdError
    from peanoclaw_amd.pool import make_description
    a = amd.register(make_description((0.0, 0.0), 1.0, 4))
    b = amd.register(make_description((0.0, 0.0), 1.0, 4))
    r = amd.register(make_description((1.0, 0.0), 1.0, 4, ny=6))
    lib = amd.lib
    assert lib.pc_apply_flux_correction(amd.h, a, b, 0, 1) == 1      # same position
    assert lib.pc_apply_flux_correction(amd.h, a, b, 2, 1) == 1      # bad dimension
    assert lib.pc_compute_fluxes(amd.h, r) == 1                      # not square
    with pytest.raises(PeanoClawAmdError):
        from peanoclaw_amd._abi import check
        check(lib.pc_apply_flux_correction(amd.h, a, r, 0, 1))
    for s in (a, b, r):
        amd.release(s)
