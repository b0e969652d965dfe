"""Golden vectors written by GRASP2018's own routines (tests/golden/make_golden.py,
from the reference's Fortran compiled into oracle/_ref).

CPU: the oracle reproduces every vector bit for bit -- grid, nuclear
potential, the GENINTRK table, RINTI, YZK and SPICMV2.
GPU: the HIP path against the same vectors: R^k keys/KSTART exact and values
within 1e-13 (relative to max(|R^k|, 1e-3 hartree)), I(a,b) within 1e-13,
SpMM within 1e-13 of the backward-error bound.
"""
import os

import numpy as np
import pytest

import oracle as O
from grasp_amd import rci

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return dict(np.load(os.path.join(HERE, name), allow_pickle=False))


class _Orb:
    def __init__(self, d):
        self.P, self.Q = d["P"], d["Q"]
        self.mf, self.nak, self.nkl, self.nkj = d["mf"], d["nak"], d["nkl"], d["nkj"]


def _oracle(d):
    g = O.Grid(int(d["n"]), float(d["rnt"]), float(d["h"]))
    o = O.Orbitals(d["P"], d["Q"], d["mf"], d["nak"], d["nkl"], d["nkj"])
    return g, o


@pytest.mark.parametrize("name", ["c3_fermi.npz", "fe17_grid.npz"])
def test_oracle_reproduces_reference_grid_and_potential(name):
    d = _load(name)
    g, o = _oracle(d)
    assert np.array_equal(g.r, d["r"]) and np.array_equal(g.rp, d["rp"])
    assert np.array_equal(g.rpor[1:], d["rpor"][1:])
    assert np.array_equal(O.nucpot(g, float(d["z"]), list(d["parm"])), d["zz"])


def test_oracle_reproduces_reference_rinti():
    for name, key in (("c3_fermi.npz", "iab"), ("fe17_grid.npz", "iaa")):
        d = _load(name)
        g, o = _oracle(d)
        nw = d["P"].shape[0]
        for a in range(nw):
            for b in range(a, nw):
                if key == "iaa" and a != b:
                    continue
                if d["nak"][a] != d["nak"][b]:
                    continue
                want = d[key][a, b] if key == "iab" else d[key][a]
                assert O.rinti(g, o, d["zz"], float(d["c"]), a + 1, b + 1, 0) == want


def test_oracle_reproduces_reference_genintrk_and_yzk():
    d = _load("c3_fermi.npz")
    g, o = _oracle(d)
    kst, keys, vals, *_ = O.genintrk(g, o)
    assert np.array_equal(kst, d["kstart"]) and np.array_equal(keys, d["keys"])
    assert np.array_equal(vals, d["vals"])
    for y, (k, a, b, mtp) in zip(d["yzk"], d["yzk_key"]):
        yo, mo = O.yzk(g, o, int(k), int(a) - 1, int(b) - 1)
        assert mo == mtp and np.array_equal(yo, y)


def test_oracle_reproduces_reference_spicmv2():
    d = _load("spicmv2.npz")
    assert np.array_equal(O.spicmv2(d["row_end"], d["col"], d["val"], d["B"]), d["C"])


@pytest.mark.gpu
def test_gpu_rk_table_and_one_body_match_reference(gpu):
    d = _load("c3_fermi.npz")
    n, h = int(d["n"]), float(d["h"])
    orb = _Orb(d)
    kst, keys, vals = rci.genintrk(d["r"], d["rp"], d["rpor"], n, h, orb)
    assert np.array_equal(kst, d["kstart"]) and np.array_equal(keys, d["keys"])
    assert np.all(np.abs(vals - d["vals"]) <= 1e-13 * np.maximum(np.abs(d["vals"]), 1e-3))
    iab = rci.one_body(d["r"], d["rp"], d["rpor"], n, h, orb, d["zz"], float(d["c"]))
    mask = d["iab"] != 0
    assert np.all(np.abs(iab[mask] - d["iab"][mask]) <= 1e-13 * np.maximum(np.abs(d["iab"][mask]), 1.0))


@pytest.mark.gpu
def test_gpu_spmm_matches_reference(gpu):
    d = _load("spicmv2.npz")
    from grasp_amd.rci import HamiltonianBlock
    hb = HamiltonianBlock.from_rows(d["row_end"], d["col"], d["val"])
    got = hb.spicmv2(d["B"])
    bound = O.spicmv2(d["row_end"], d["col"], np.abs(d["val"]), np.abs(d["B"]))
    assert np.all(np.abs(got - d["C"]) <= 1e-13 * bound + 1e-300)
