"""Writes tests/golden/he_3pn6_lists.npz: the angular-coefficient lists of the
reference-held He CSF list (src/appl/rhfszeeman95/examples/data_He/3Pn6.c:
36 subshells up to 6h, blocks of 70, 180 and 222 CSFs) as the reference's
own librang90 produces them for rci90's element loop (ONESCALAR and
RKCO_GG per pair IR <= IC, rci90/setham_gg.f90:213-292), with SETHAM's
transverse Breit lists (RKCO_GG with BREID, setham_gg.f90:298-302: six
labels a, b, c, d, k and the integral type, negative for core terms).

oracle/build_rang.sh compiles the producer from the reference sources into
oracle/_ref/rang_lists; this script runs it on each block and stores the
lists in the layout of gci_pair_lists (include/graspci.h), with the
subshells' quantum numbers, so GPU tests can assemble H from them.

Run:  python tests/golden/make_rang_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "rang_lists")
CSF = os.path.join(os.environ.get("GRASP_REFERENCE", "/root/reference"),
                   "src/appl/rhfszeeman95/examples/data_He/3Pn6.c")
OUT = os.path.join(HERE, "he_3pn6_lists.npz")


def parse(path):
    """The driver's stream (see oracle/rang_driver.f90) as gci_pair_lists arrays."""
    b = open(path, "rb").read()
    nw, ncf, ncore, breit = (int(x) for x in np.frombuffer(b, dtype=np.int32, count=4))
    off = 16
    q = np.frombuffer(b, dtype=np.int32, count=4 * nw, offset=off).reshape(nw, 4).copy()
    off += 16 * nw
    ic, ir, ia, ib, tsh_off, tb_beg, tb_end = ([] for _ in range(7))
    tsh, lab, coe = [], [], []
    bt_beg, bt_end, blab, bcoe = [], [], [], []
    rec_t = np.dtype([("l", "<i4", 5), ("c", "<f8")])
    brec_t = np.dtype([("l", "<i4", 6), ("c", "<f8")])
    while off < len(b):
        c_ic, c_ir, c_ia, c_ib, nt = (int(x) for x in np.frombuffer(b, dtype=np.int32, count=5, offset=off))
        off += 20
        ic.append(c_ic)
        ir.append(c_ir)
        ia.append(c_ia)
        ib.append(c_ib)
        tsh_off.append(len(tsh))
        if nt:
            tsh.extend(np.frombuffer(b, dtype=np.float64, count=nt, offset=off))
            off += 8 * nt
        nv = int(np.frombuffer(b, dtype=np.int32, count=1, offset=off)[0])
        off += 4
        tb_beg.append(len(coe))
        if nv:
            rec = np.frombuffer(b, dtype=rec_t, count=nv, offset=off)
            off += 28 * nv
            lab.extend(rec["l"])
            coe.extend(rec["c"])
        tb_end.append(len(coe))
        if breit:
            nb = int(np.frombuffer(b, dtype=np.int32, count=1, offset=off)[0])
            off += 4
            bt_beg.append(len(bcoe))
            if nb:
                rec = np.frombuffer(b, dtype=brec_t, count=nb, offset=off)
                off += 32 * nb
                blab.extend(rec["l"])
                bcoe.extend(rec["c"])
            bt_end.append(len(bcoe))
    assert off == len(b)
    return dict(nw=nw, ncf=ncf, ncore=ncore, np=q[:, 0], nak=q[:, 1], nkl=q[:, 2], nkj=q[:, 3],
                pair_ic=np.array(ic, np.int32), pair_ir=np.array(ir, np.int32),
                one_ia=np.array(ia, np.int32), one_ib=np.array(ib, np.int32),
                tsh_off=np.array(tsh_off, np.int64), tshell=np.array(tsh, np.float64),
                tb_beg=np.array(tb_beg, np.int64), tb_end=np.array(tb_end, np.int64),
                label=np.array(lab, np.int32).reshape(-1), coeff=np.array(coe, np.float64),
                bt_beg=np.array(bt_beg, np.int64), bt_end=np.array(bt_end, np.int64),
                blabel=np.array(blab, np.int32).reshape(-1), bcoeff=np.array(bcoe, np.float64))


def main():
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} missing: run oracle/build_rang.sh")
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for blk in (1, 2, 3):
            path = os.path.join(td, f"b{blk}.bin")
            subprocess.run([EXE, CSF, str(blk), path, "1"], check=True, stdout=subprocess.DEVNULL)
            d = parse(path)
            print(f"block {blk}: ncf={d['ncf']} pairs={len(d['pair_ic'])} tshell={len(d['tshell'])} "
                  f"two-body={len(d['coeff'])} breit={len(d['bcoeff'])}")
            for k, v in d.items():
                out[f"b{blk}_{k}"] = np.asarray(v)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
