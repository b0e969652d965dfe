"""H assembly (SETHAM element loop, rci90/setham_gg.f90:162-385) on the GPU
vs the oracle.  Given the same integral tables the GPU sums each pair's
terms in list order with contraction off, so elements, structure and EAV
must be bit-identical."""
import numpy as np
import pytest

import oracle as O
from grasp_amd import _lib as L
from grasp_amd import rci
from grasp_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables():
    z = 6.0
    n, h = 590, 0.05
    g = O.Grid(n, 2e-6 / z, h)
    orb = W.screened_orbitals(z, W.subshells(4, 3), occupied={1: 2, 2: 2})
    o = O.Orbitals(orb.P, orb.Q, orb.mf, orb.nak, orb.nkl, orb.nkj)
    kst, keys, vals, kv, labels = O.genintrk(g, o)
    zz = O.nucpot(g, z, W.fermi_parameters(12.0))
    nw = o.nw
    iab = np.zeros((nw, nw))
    for a in range(1, nw + 1):
        for b in range(a, nw + 1):
            if orb.nak[a - 1] == orb.nak[b - 1]:
                iab[a - 1, b - 1] = O.rinti(g, o, zz, W.C_LIGHT, a, b, 0)
    return dict(g=g, o=o, orb=orb, kst=kst, keys=keys, vals=vals, kv=kv, labels=labels, iab=iab, nw=nw)


@pytest.mark.parametrize("ncf,fill,seed", [(1, 0.0, 1), (57, 0.3, 2), (400, 0.1, 3), (1500, 0.05, 4)])
def test_setham_bitexact(gpu, tables, ncf, fill, seed):
    T = tables
    pattern = W.clustered_block(ncf, seed, fill=fill)
    lists = W.synthetic_lists(pattern, T["nw"], T["orb"].nkl, T["labels"], T["kv"], seed)
    row_end, col, val, eav_sum = O.setham(ncf, T["nw"], lists, T["iab"], T["kst"], T["keys"], T["vals"])
    blk, eav, nel = rci.setham(ncf, T["nw"], lists, T["iab"], T["kst"], T["keys"], T["vals"])
    assert nel == len(col)
    assert eav == eav_sum / ncf
    iendc, irow, emt = blk.download()
    assert np.array_equal(iendc, row_end)
    assert np.array_equal(irow, col)
    want = val.copy()
    want[row_end[1:] - 1] = want[row_end[1:] - 1] - eav
    assert np.array_equal(emt, want)


def test_setham_lfordr(gpu, tables):
    T = tables
    ncf = 300
    pattern = W.clustered_block(ncf, 7, fill=0.2)
    lists = W.synthetic_lists(pattern, T["nw"], T["orb"].nkl, T["labels"], T["kv"], 7)
    row_end, col, val, eav_sum = O.setham(ncf, T["nw"], lists, T["iab"], T["kst"], T["keys"], T["vals"],
                                          lfordr=True, iccut=40)
    blk, eav, nel = rci.setham(ncf, T["nw"], lists, T["iab"], T["kst"], T["keys"], T["vals"], lfordr=True,
                               iccut=40)
    iendc, irow, emt = blk.download()
    assert np.array_equal(iendc, row_end) and np.array_equal(irow, col)
    off = np.ones(len(col), bool)
    off[row_end[1:] - 1] = False
    assert np.all(col[off] <= 40)


def test_setham_missing_integral(gpu, tables):
    T = tables
    pattern = W.clustered_block(20, 9, fill=0.3)
    lists = W.synthetic_lists(pattern, T["nw"], T["orb"].nkl, T["labels"], T["kv"], 9)
    lists["label"] = lists["label"].copy()
    lists["label"][4] = 19  # k beyond the table
    lists["coeff"] = lists["coeff"].copy()
    lists["coeff"][0] = 0.5
    with pytest.raises(L.GciError) as e:
        rci.setham(20, T["nw"], lists, T["iab"], T["kst"], T["keys"], T["vals"])
    assert e.value.code == -4


def test_setham_split_then_solve(gpu, tables):
    """Assembly into the split form feeds the solver directly."""
    T = tables
    ncf = 2200
    pattern = W.clustered_block(ncf, 12, fill=0.04)
    lists = W.synthetic_lists(pattern, T["nw"], T["orb"].nkl, T["labels"], T["kv"], 12)
    blk, eav, nel = rci.setham(ncf, T["nw"], lists, T["iab"], T["kst"], T["keys"], T["vals"],
                               precision=L.GCI_SPLIT32)
    iendc, irow, emt = blk.download()
    B = np.asfortranarray(np.random.default_rng(0).standard_normal((ncf, 3)))
    got = blk.spicmv2(B, L.GCI_OP_SPLIT_FULL)
    want = O.spicmv2(iendc, irow, emt, B)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
