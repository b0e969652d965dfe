"""R^k tables (GENINTRK/SLATER/YZK/QUAD) and one-body I(ab) (RINTI) on the
GPU vs the oracle.  Keys, packing and KSTART must match exactly; values
within 1e-13 of the oracle relative to max(|R^k|, 1e-3 hartree) -- the GPU
evaluates the Newton-Cotes recurrences as prefix scans and the quadrature as
a tree sum, so only the summation order differs."""
import numpy as np
import pytest

import oracle as O
from grasp_amd import rci
from grasp_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = {
    # C III (Z=6): 1s..4f, 16 orbitals, Fermi nucleus
    "C_III": (6.0, 4, 3, {1: 2, 2: 2}, 12.0),
    # Fe XVII (Z=26): n <= 5, l <= 4 (25 orbitals)
    "Fe_XVII": (26.0, 5, 4, {1: 2, 2: 8}, 56.0),
}


def _setup(z, nmax, lmax, occ):
    n, h = 590, 0.05
    rnt = 2e-6 / z
    orb = W.screened_orbitals(z, W.subshells(nmax, lmax), occupied=occ)
    r, rp, rpor = rci.radgrd(n, rnt, h)
    g = O.Grid(n, rnt, h)
    assert np.array_equal(r, g.r) and np.array_equal(rp, g.rp)
    o = O.Orbitals(orb.P, orb.Q, orb.mf, orb.nak, orb.nkl, orb.nkj)
    return n, h, r, rp, rpor, g, o, orb


@pytest.mark.parametrize("case", list(CASES))
def test_rk_table_matches_oracle(gpu, case):
    z, nmax, lmax, occ, amass = CASES[case]
    n, h, r, rp, rpor, g, o, orb = _setup(z, nmax, lmax, occ)
    kst, keys, vals = rci.genintrk(r, rp, rpor, n, h, orb)
    okst, okeys, ovals, kv, labels = O.genintrk(g, o)
    assert np.array_equal(kst, okst)
    assert np.array_equal(keys, okeys)
    err = np.abs(vals - ovals)
    scale = np.maximum(np.abs(ovals), 1e-3)
    rel = err / scale
    print(f"{case}: {len(keys)} integrals, max rel err {rel.max():.2e}, bit-exact {np.mean(err == 0):.3f}")
    assert np.all(rel <= 1e-13)


@pytest.mark.parametrize("case", list(CASES))
def test_one_body_matches_oracle(gpu, case):
    z, nmax, lmax, occ, amass = CASES[case]
    n, h, r, rp, rpor, g, o, orb = _setup(z, nmax, lmax, occ)
    parm = W.fermi_parameters(amass)
    zz = O.nucpot(g, z, parm)
    iab = rci.one_body(r, rp, rpor, n, h, orb, zz, W.C_LIGHT)
    nw = o.nw
    for a in range(1, nw + 1):
        for b in range(a, nw + 1):
            if orb.nak[a - 1] != orb.nak[b - 1]:
                assert iab[a - 1, b - 1] == 0.0
                continue
            want = O.rinti(g, o, zz, W.C_LIGHT, a, b, 0)
            assert abs(iab[a - 1, b - 1] - want) <= 1e-13 * max(abs(want), 1e-3), (a, b)


def test_rk_table_rejects_bad_orbitals(gpu):
    n, h, r, rp, rpor, g, o, orb = _setup(6.0, 2, 1, {1: 2})
    orb.mf[0] = n  # beyond N-4
    from grasp_amd._lib import GciError
    with pytest.raises(GciError):
        rci.genintrk(r, rp, rpor, n, h, orb)
