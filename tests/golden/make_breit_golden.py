"""Writes tests/golden/he_3pn6_breit.npz: the oracle's block Hamiltonians of
the reference-held He CSF list (data_He/3Pn6.c, the lists of
he_3pn6_lists.npz) with the transverse Breit terms, assembled the way rci90
does a run: GENINTBREIT1 and GENINTBREIT2 once, then the three blocks in
order through one Breit state (BESSEL's cache and the BRINT3..6 values carry
over from block to block, rci90/rci92.f90:198-202, setham_gg.f90:298-357).

Orbitals: screened-hydrogenic estimates for Z = 2 on the N = 590 grid
(grasp_amd.workloads), orbital energies E = their binding energies,
WFACT = 1, c = 137.036 (the speed of light workloads.C_LIGHT).  The oracle
is pinned bit for bit against the reference build (tests/test_ref_pins.py).

Stored per block b: b{b}_row_end, b{b}_col, b{b}_val (SETHAM's rows,
diagonal last, unshifted), b{b}_eav_sum (diagonal sum), b{b}_elsto.
The full tables take the oracle about a minute; run once:
    python tests/golden/make_breit_golden.py
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from grasp_amd import workloads as W  # noqa: E402

LISTS = os.path.join(HERE, "he_3pn6_lists.npz")
OUT = os.path.join(HERE, "he_3pn6_breit.npz")
KEYS = ("pair_ic", "pair_ir", "one_ia", "one_ib", "tsh_off", "tshell", "tb_beg", "tb_end", "label", "coeff")
BKEYS = ("bt_beg", "bt_end", "blabel", "bcoeff")


def he_setup():
    z, n, h = 2.0, 590, 0.05
    rnt = 2e-6 / z
    orb = W.screened_orbitals(z, W.subshells(6, 5))
    g = O.Grid(n, rnt, h)
    o = O.Orbitals(orb.P, orb.Q, orb.mf, orb.nak, orb.nkl, orb.nkj)
    return z, g, o, orb


def main():
    G = np.load(LISTS)
    z, g, o, orb = he_setup()
    t0 = time.time()
    kst, keys, vals, kv, labels = O.genintrk(g, o)
    zz = O.nucpot(g, z)
    nw = o.nw
    iab = np.zeros((nw, nw))
    for a in range(1, nw + 1):
        for b in range(a, nw + 1):
            if orb.nak[a - 1] == orb.nak[b - 1]:
                iab[a - 1, b - 1] = O.rinti(g, o, zz, W.C_LIGHT, a, b, 0)
    br = O.Breit(g, o, orb.ebind, W.C_LIGHT, 1.0)
    n1 = len(br.genintbreit(1)[1])
    n2 = len(br.genintbreit(2)[1])
    print(f"tables: {n1} type 1, {n2} type 2 in {time.time() - t0:.1f}s")
    out = {}
    for b in (1, 2, 3):
        ncf = int(G[f"b{b}_ncf"])
        L = {k: G[f"b{b}_{k}"] for k in KEYS}
        B = {k: G[f"b{b}_{k}"] for k in BKEYS}
        row_end, col, val, eav_sum, elsto = O.setham_tb(ncf, nw, L, B, br, iab, kst, keys, vals)
        out[f"b{b}_row_end"] = row_end
        out[f"b{b}_col"] = col
        out[f"b{b}_val"] = val
        out[f"b{b}_eav_sum"] = np.float64(eav_sum)
        out[f"b{b}_elsto"] = np.float64(elsto)
        print(f"block {b}: ncf={ncf} nelmnt={len(col)} eav={eav_sum / ncf + elsto:.12f} elsto={elsto}")
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes) in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    main()
