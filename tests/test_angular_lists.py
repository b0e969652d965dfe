"""The angular-coefficient list fixtures (tests/golden/he_3pn6_lists.npz,
written by the reference's own librang90 through oracle/_ref/rang_lists from
the reference-held He CSF list data_He/3Pn6.c) against what the physics
fixes, and the oracle's SETHAM on them.

* Pair counts: NCF (NCF+1)/2 per block, the lower-triangle element counts
  of the reference's 3Pn6.csum (2485, 16290, 24753).
* One-body: for every diagonal pair the coefficients of the scalar one-body
  operator are the subshell occupations, so they sum to the number of
  electrons (2) -- ONESCALAR, librang90/onescalar.f90.
* Two-body: for every diagonal pair the direct monopole coefficients
  (a b a b, k = 0) sum to the number of electron pairs N(N-1)/2 = 1 --
  RKCO_GG, librang90/rkco_gg.f90.
* Labels index the 36 subshells in the .c file's order (1s ... 6h), which is
  grasp_amd.workloads.subshells(6, 5); k parity and triangle rules hold.
"""
import os

import numpy as np
import pytest

from grasp_amd import workloads as W

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "he_3pn6_lists.npz"))
BLOCKS = {1: 70, 2: 180, 3: 222}


def lists(b):
    keys = ("pair_ic", "pair_ir", "one_ia", "one_ib", "tsh_off", "tshell", "tb_beg", "tb_end", "label", "coeff")
    return {k: G[f"b{b}_{k}"] for k in keys}


@pytest.mark.parametrize("b", list(BLOCKS))
def test_pair_counts_and_subshells(b):
    ncf = int(G[f"b{b}_ncf"])
    assert ncf == BLOCKS[b]
    L = lists(b)
    assert len(L["pair_ic"]) == ncf * (ncf + 1) // 2
    # pairs in SETHAM's order: IC ascending, IR = 1..IC
    ic, ir = L["pair_ic"], L["pair_ir"]
    want_ic = np.repeat(np.arange(1, ncf + 1), np.arange(1, ncf + 1))
    assert np.array_equal(ic, want_ic)
    assert np.all((ir >= 1) & (ir <= ic))
    names = W.subshells(6, 5)
    for nm, n, l, k in zip(names, G[f"b{b}_np"], G[f"b{b}_nkl"], G[f"b{b}_nak"]):
        assert W.parse_subshell(nm) == (int(n), int(l), int(k))


@pytest.mark.parametrize("b", list(BLOCKS))
def test_diagonal_coefficients_count_electrons_and_pairs(b):
    L = lists(b)
    nw = int(G[f"b{b}_nw"])
    ic, ir = L["pair_ic"], L["pair_ir"]
    lab = L["label"].reshape(-1, 5)
    for p in np.nonzero(ic == ir)[0]:
        assert L["one_ia"][p] != 0 and L["one_ia"][p] == L["one_ib"][p]
        t = L["tshell"][L["tsh_off"][p]: L["tsh_off"][p] + nw]
        assert abs(t.sum() - 2.0) < 1e-14
        s = slice(L["tb_beg"][p], L["tb_end"][p])
        lb, c = lab[s], L["coeff"][s]
        direct0 = (lb[:, 4] == 0) & (lb[:, 0] == lb[:, 2]) & (lb[:, 1] == lb[:, 3])
        assert abs(c[direct0].sum() - 1.0) < 1e-14


@pytest.mark.parametrize("b", list(BLOCKS))
def test_two_body_labels_obey_selection_rules(b):
    L = lists(b)
    lab = L["label"].reshape(-1, 5)
    nkl = G[f"b{b}_nkl"]
    nkj = G[f"b{b}_nkj"]
    a, bb, c, d, k = (lab[:, i] for i in range(5))
    assert np.all((a >= 1) & (a <= 36) & (bb >= 1) & (c >= 1) & (d >= 1) & (k >= 0))
    la, lb_, lc, ld = nkl[a - 1], nkl[bb - 1], nkl[c - 1], nkl[d - 1]
    # parity: l_a + k + l_c and l_b + k + l_d even
    assert np.all((la + k + lc) % 2 == 0) and np.all((lb_ + k + ld) % 2 == 0)
    # triangle on j: |j_a - j_c| <= k <= j_a + j_c (2j stored)
    ja, jc = nkj[a - 1], nkj[c - 1]
    assert np.all(2 * k >= np.abs(ja - jc)) and np.all(2 * k <= ja + jc)


# Lowest levels of each block held by the reference
# (src/appl/rhfszeeman95/examples/data_He/3Pn6.csum): J=0-, J=1- (two), J=2-.
# They come from MCDHF-optimised orbitals with the transverse Breit
# interaction, so they pin this Dirac-Coulomb CI on He+ hydrogenic orbitals
# only loosely (a few mhartree).
CSUM_LEVELS = {1: [-2.13325863106008], 2: [-2.13326313263103, -2.12392844472579], 3: [-2.13326348364110]}


def he_tables():
    import oracle as O

    z, n, h = 2.0, 590, 0.05
    names = W.subshells(6, 5)
    orb = W.screened_orbitals(z, names)  # He+ hydrogenic: orthonormal within each kappa
    g = O.Grid(n, 2e-6 / z, h)
    o = O.Orbitals(orb.P, orb.Q, orb.mf, orb.nak, orb.nkl, orb.nkj)
    kst, keys, vals, kv, labels = O.genintrk(g, o)
    zz = O.nucpot(g, z)
    nw = o.nw
    iab = np.zeros((nw, nw))
    for a in range(1, nw + 1):
        for bb in range(a, nw + 1):
            if orb.nak[a - 1] == orb.nak[bb - 1]:
                iab[a - 1, bb - 1] = O.rinti(g, o, zz, W.C_LIGHT, a, bb, 0)
    return dict(z=z, n=n, h=h, names=names, orb=orb, g=g, o=o, zz=zz, kst=kst, keys=keys, vals=vals, iab=iab)


@pytest.mark.parametrize("b", list(BLOCKS))
def test_oracle_ci_on_reference_lists_matches_reference_levels(b):
    """The oracle's SETHAM element loop and MANEIG on the reference's lists:
    rows end on their diagonal, and the lowest levels agree with the
    reference-held energies of the same CSF blocks to 5 mhartree."""
    import oracle as O

    T = he_tables()
    ncf = int(G[f"b{b}_ncf"])
    row_end, col, val, eav_sum = O.setham(ncf, T["o"].nw, lists(b), T["iab"], T["kst"], T["keys"], T["vals"])
    assert np.array_equal(col[row_end[1:] - 1], np.arange(1, ncf + 1))
    eav = eav_sum / ncf
    want = np.array(CSUM_LEVELS[b])
    w, *_ = O.maneig(row_end, col, val, eav, len(want))
    assert np.all(np.abs(w + eav - want) < 5e-3), (w + eav, want)
