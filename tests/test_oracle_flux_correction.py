"""The oracle's flux correction against a cell-by-cell brute force.

DefaultFluxCorrection (interSubgridCommunication/DefaultFluxCorrection.cpph:
13-164), its face iterator (grid/SubgridFaceIterator.cpph) and
HyperplaneRegion (geometry/HyperplaneRegion.cpp:25-103), as the present
code behaves, restated independently here from what each routine computes:

computeFluxes: per face cell, 0.5 * (F(ghost uOld) + F(interior uOld)) with
  Numerics::computeFlux (Numerics.h:185-203): F_i = u_i * u_{1+d}, the
  momenta times h once more, + 0.5 * 1.0 * h * h on the normal momentum;
  the upwind term of the statement at :70-72 is discarded by its ';'.
applyCorrection: time overlap of the two intervals; for every source cell
  of the face (dim, dir), the destination cells whose tangential index lies
  in [floor((lo - p + eps) / s), ceil((hi - p - eps) / s)) where the interval
  is the source cell's clipped to the destination's FIRST subcell
  (getSubcellSize() used as the patch size); the destination flux is read
  from the start of the face whatever that index is; uNew of the interior
  cell next to the face gains (F_src - F_dst) * overlap time * overlap area
  / subcell area.

Cases: 2 x 2 and 3 x 3 coarse patches with a fine neighbour (factor 3) on
each of the four sides, both roles (coarse source, fine source), random
data and times.
"""
import math

import numpy as np
import pytest

from oracle_lib import Patch, dptr, lib


def face_offsets(nx, ny, nq=3):
    o = [0, ny * nq, 2 * ny * nq, 2 * ny * nq + nx * nq]
    return {(0, -1): o[0], (0, 1): o[1], (1, -1): o[2], (1, 1): o[3]}


def ref_flux(dim, u):
    f = [u[i] * u[1 + dim] for i in range(3)]
    f[1] *= u[0]
    f[2] *= u[0]
    f[1 + dim] += 0.5 * 1.0 * u[0] * u[0]
    return f


def brute_fluxes(P):
    nx, ny = P.s.nx, P.s.ny
    fo = face_offsets(nx, ny)
    out = np.zeros(2 * 3 * (nx + ny))
    uo = np.array(P.uold)  # [q, x + 1, y + 1]
    for dim in (0, 1):
        nf = ny if dim == 0 else nx
        for d in (-1, 1):
            # interior / ghost normal index (ny for an upper face, as the
            # iterator reads getSubdivisionFactor(+1); square patches here)
            ni, ng = (0, -1) if d == -1 else (ny - 1, ny)
            for k in range(nf):
                ci = (ni, k) if dim == 0 else (k, ni)
                cg = (ng, k) if dim == 0 else (k, ng)
                inner = [uo[q, ci[0] + 1, ci[1] + 1] for q in range(3)]
                outer = [uo[q, cg[0] + 1, cg[1] + 1] for q in range(3)]
                fi, fg = ref_flux(dim, inner), ref_flux(dim, outer)
                for q in range(3):
                    out[fo[(dim, d)] + q * nf + k] = 0.5 * (fg[q] + fi[q])
    return out


def brute_apply(S, sflux, D, dflux, dim, d):
    sp, dp = list(S.s.pos), list(D.s.pos)
    ss, ds = list(S.s.subcell), list(D.s.subcell)
    t_ov = min(S.s.time + S.s.dt, D.s.time + D.s.dt) - max(S.s.time, D.s.time)
    rdim = None
    for k in (0, 1):
        if abs(sp[k] - dp[k]) > 1e-8:
            rdim = k
    sd = 1 - rdim
    od = 1 - dim
    out = np.array(D.unew)
    nfs = S.s.ny if dim == 0 else S.s.nx
    fos = face_offsets(S.s.nx, S.s.ny)[(dim, d)]
    fod = face_offsets(D.s.nx, D.s.ny)[(dim, -d)]
    nfd = D.s.ny if dim == 0 else D.s.nx
    ni = 0 if -d == -1 else D.s.ny - 1
    for ks in range(nfs):
        s_lo = sp[sd] + ss[sd] * ks
        lo = max(dp[sd], s_lo)
        hi = min(dp[sd] + ds[sd], s_lo + ss[sd])
        first = int((lo - dp[sd] + 1e-10) / ds[sd])
        last = math.ceil((hi - dp[sd] - 1e-10) / ds[sd])
        for j, t in enumerate(range(first, last)):
            a_lo, a_hi = sp[od] + ks * ss[od], sp[od] + (ks + 1) * ss[od]
            b_lo, b_hi = dp[od] + t * ds[od], dp[od] + (t + 1) * ds[od]
            area = max(0.0, min(a_hi, b_hi) - max(a_lo, b_lo))
            cell = (ni, t) if dim == 0 else (t, ni)
            for q in range(3):
                out[q][cell] += (sflux[fos + q * nfs + ks] - dflux[fod + q * nfd + j]) * \
                    t_ov * area / (ds[0] * ds[1])
    return out


def _patch(rng, n, pos, size, level, time, dt):
    P = Patch(3, 1, n, 1, pos, size, level, time, dt)
    P.unew[0] = rng.uniform(0.5, 2.0, (n, n))
    P.unew[1:] = rng.uniform(-0.4, 0.4, (2, n, n))
    P.uold[0] = rng.uniform(0.5, 2.0, (n + 2, n + 2))
    P.uold[1:] = rng.uniform(-0.4, 0.4, (2, n + 2, n + 2))
    return P


def test_compute_fluxes_brute_force():
    rng = np.random.default_rng(1)
    for n in (2, 3, 5):
        P = _patch(rng, n, (0.25, 0.5), 0.3, 0, 0.1, 0.01)
        f = np.zeros(lib().or_flux_array_size(P.ref))
        lib().or_compute_fluxes(P.ref, dptr(f))
        assert f.size == 12 * n
        np.testing.assert_allclose(f, brute_fluxes(P), rtol=1e-15, atol=1e-16)


def _pairs():
    """(coarse n, side, fine index along the face): a coarse patch [1,2]^2
    and a fine neighbour of a third of its size across `side`."""
    for n in (2, 3):
        for dim in (0, 1):
            for d in (-1, 1):
                for k in range(3):
                    yield n, dim, d, k


@pytest.mark.parametrize("n,dim,d,k", list(_pairs()))
def test_apply_correction_brute_force(n, dim, d, k):
    rng = np.random.default_rng(100 * n + 10 * dim + 3 * (d + 1) + k)
    cpos = (1.0, 1.0)
    fs = 1.0 / 3.0
    # fine patch across side (dim, d) of the coarse patch, k-th along it
    fpos = [1.0 + k * fs, 1.0 + k * fs]
    fpos[dim] = 1.0 - fs if d == -1 else 2.0
    C = _patch(rng, n, cpos, 1.0, 0, 0.2, 0.05)
    F = _patch(rng, n, tuple(fpos), fs, 1, 0.21, 0.02)
    fc = rng.uniform(-1, 1, 12 * n)
    ff = rng.uniform(-1, 1, 12 * n)
    for S, sf, D, df, dd in ((C, fc, F, ff, d), (F, ff, C, fc, -d)):
        want = brute_apply(S, sf, D, df, dim, dd)
        rc = lib().or_apply_flux_correction(S.ref, dptr(sf), D.ref, dptr(df), dim, dd)
        assert rc == 0
        np.testing.assert_allclose(np.array(D.unew), want, rtol=1e-14, atol=1e-15)


def test_apply_correction_reaches_only_the_first_tangential_subcell():
    """getInterfaceRegionForSubcell takes the destination's subcell size as
    the patch size (HyperplaneRegion.cpp:76): a coarse source corrects only
    the fine destination cell at tangential index 0, and a fine source only
    the coarse cell at index 0 (when it overlaps it)."""
    rng = np.random.default_rng(5)
    C = _patch(rng, 3, (1.0, 1.0), 1.0, 0, 0.2, 0.05)
    F = _patch(rng, 3, (2.0, 1.0), 1.0 / 3, 1, 0.2, 0.05)  # right of C, first third
    fc, ff = rng.uniform(-1, 1, 36), rng.uniform(-1, 1, 36)
    before = np.array(F.unew)
    assert lib().or_apply_flux_correction(C.ref, dptr(fc), F.ref, dptr(ff), 0, 1) == 0
    changed = np.argwhere((np.array(F.unew) != before).any(axis=0))
    assert changed.tolist() == [[0, 0]]
    before = np.array(C.unew)
    assert lib().or_apply_flux_correction(F.ref, dptr(ff), C.ref, dptr(fc), 0, -1) == 0
    changed = np.argwhere((np.array(C.unew) != before).any(axis=0))
    assert changed.tolist() == [[2, 0]]


def test_apply_correction_equal_positions_is_rejected():
    rng = np.random.default_rng(6)
    A = _patch(rng, 2, (0.0, 0.0), 1.0, 0, 0.0, 0.1)
    B = _patch(rng, 2, (0.0, 0.0), 1.0, 0, 0.0, 0.1)
    f = np.zeros(24)
    assert lib().or_apply_flux_correction(A.ref, dptr(f), B.ref, dptr(f), 0, 1) == -1
