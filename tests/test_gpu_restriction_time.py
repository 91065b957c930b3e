"""The restriction's [0, 1] time basis on the GPU (through the C ABI).

A fine leaf over [t0, t0 + dt] (dt > 0, t0 > 0) holding data linear in
time is restricted into its virtual parent (SWEKernel::restrictSolution +
postProcessRestriction, SWEKernel.cpp:288-316; DefaultRestriction.cpph:35-52)
and the virtual patch's column next to a coarse neighbour is transferred
into the neighbour's ghost column at the neighbour's time T
(DefaultTransfer.cpph:30-41).  Every array is compared bitwise with the
oracle, and the ghost column with the closed form: the area average of
the fine data at time T (h read as h + b on the fine side, the virtual
patch's own b subtracted afterwards).
"""
import numpy as np
import pytest

from oracle_lib import Patch as OPatch, iarr, lib as olib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import peanoclaw_amd  # noqa: F401
    from peanoclaw_amd.pool import PatchPool
    pool = PatchPool(0)
    yield pool
    pool.close()


def _opatch(unew, uold, b, pos, size, level, time, dt):
    n = unew.shape[1]
    p = OPatch(3, 1, n, 1, pos, size, level, time, dt)
    p.unew[:] = unew
    p.uold[:] = uold
    p.pwg[0] = b
    return p


@pytest.mark.parametrize("t0,dt", [(0.37, 0.05), (2.5, 0.2)])
def test_restriction_time_basis_bitwise_and_closed_form(amd, t0, dt):
    from peanoclaw_amd.numerics import Patch, SWENumerics
    from peanoclaw_amd.pool import make_description
    rng = np.random.default_rng(int(t0 * 100))
    num = SWENumerics(amd)
    nf, nc = 9, 3
    fpos = (2.0 / 3, 1.0 / 3)  # covers the virtual cell (2, 1)
    f = np.stack([rng.uniform(0.8, 1.6, (nf, nf)), rng.uniform(-0.3, 0.3, (nf, nf)),
                  rng.uniform(-0.3, 0.3, (nf, nf))])
    g = rng.uniform(-2.0, 2.0, (3, nf, nf))
    fold = np.zeros((3, nf + 2, nf + 2))
    fold[:, 1:-1, 1:-1] = f
    fnew = f + dt * g
    fb = rng.uniform(0.0, 0.3, (nf + 2, nf + 2))
    vb = rng.uniform(0.0, 0.3, (nc + 2, nc + 2))
    sf = amd.register(make_description(fpos, 1.0 / 3, nf, level=1, time=t0, timestep_size=dt))
    sv = amd.register(make_description((0.0, 0.0), 1.0, nc, time=0.0, timestep_size=0.0,
                                       is_virtual=True))
    amd.upload(sf, unew=fnew, uold=fold, b=fb)
    amd.upload(sv, unew=np.zeros((3, nc, nc)), uold=np.zeros((3, nc + 2, nc + 2)), b=vb)
    of = _opatch(fnew, fold, fb, fpos, 1.0 / 3, 1, t0, dt)
    ov = _opatch(np.zeros((3, nc, nc)), np.zeros((3, nc + 2, nc + 2)), vb, (0.0, 0.0), 1.0,
                 0, 0.0, 0.0)
    ov.s.is_virtual = 1
    ov.s.is_leaf = 0
    print(f"restrict 9^2 leaf over [{t0}, {t0 + dt}] into a virtual 3^2 patch", flush=True)
    num.restrictSolution(Patch(amd, sf), Patch(amd, sv), False)
    olib().or_restrict_swe(of.ref, ov.ref, 0)
    num.postProcessRestriction(Patch(amd, sv), False)
    olib().or_post_process_restriction_swe(ov.ref, 0)
    gn, go, _ = amd.download(sv)
    assert np.array_equal(gn, ov.unew) and np.array_equal(go, ov.uold)
    gn, go, _ = amd.download(sf)
    assert np.array_equal(gn, of.unew) and np.array_equal(go, of.uold)
    avg_hb = (f[0] + fb[1:-1, 1:-1]).mean()
    for T in (t0 + 0.5 * dt, t0 + dt, t0 + 3.0 * dt):
        nb_u = rng.uniform(0.8, 1.6, (3, nc, nc))
        nb_o = rng.uniform(0.8, 1.6, (3, nc + 2, nc + 2))
        nb_b = rng.uniform(0.0, 0.3, (nc + 2, nc + 2))
        sn = amd.register(make_description((1.0, 0.0), 1.0, nc, time=T - 0.01,
                                           timestep_size=0.01))
        amd.upload(sn, unew=nb_u, uold=nb_o, b=nb_b)
        on = _opatch(nb_u, nb_o, nb_b, (1.0, 0.0), 1.0, 0, T - 0.01, 0.01)
        # the virtual patch's last column -> the neighbour's lower ghost column
        # (FillGhostlayerFaceFunctor, GhostlayerCompositorFunctors.cpp:37-59)
        num.transferGhostlayer((1, nc), (nc - 1, 0), (-1, 0), Patch(amd, sv), Patch(amd, sn))
        olib().or_transfer_ghostlayer(iarr(1, nc), iarr(nc - 1, 0), iarr(-1, 0), ov.ref, on.ref)
        gn, go, _ = amd.download(sn)
        assert np.array_equal(gn, on.unew) and np.array_equal(go, on.uold)
        ghost = go[:, 0, 2]  # ghost cell (-1, 1) <- virtual cell (2, 1)
        want_h = avg_hb + (T - t0) * g[0].mean() - vb[3, 2]
        want_hu = f[1].mean() + (T - t0) * g[1].mean()
        want_hv = f[2].mean() + (T - t0) * g[2].mean()
        tol = 1e-12 * (1.0 + t0 / dt)
        assert ghost[0] == pytest.approx(want_h, abs=tol)
        assert ghost[1] == pytest.approx(want_hu, abs=tol)
        assert ghost[2] == pytest.approx(want_hv, abs=tol)
        amd.release(sn)
    amd.release(sf)
    amd.release(sv)
