"""Writes the golden vectors in this directory from GRASP2018's own routines.

Run in a tree where the reference sources are present and oracle/_ref has
been built (oracle/build_ref.sh compiles RADGRD, NUCPOT/ES, YZK, SLATER,
RINTI, GENINTRK and SPICMV2 from the reference's Fortran with amdflang):

    python tests/golden/make_golden.py

Inputs are fixed (screened-hydrogenic orbitals from grasp_amd.workloads, a
seeded clustered block) and stored beside the outputs, so the fixtures are
self-contained data: the tests need neither the reference nor this script.

  c3_fermi.npz   C III (Z=6), Fermi nucleus, 1s..4f (16 orbitals), N=590:
                 grid r/rp/rpor, nuclear potential zz, the GENINTRK table
                 (kstart, keys, vals), RINTI I(a,b) for every same-kappa
                 pair a <= b, YZK for a few (k, a, b)
  fe17_grid.npz  Fe XVII (Z=26), Fermi nucleus, n <= 3: grid, zz and the
                 diagonal RINTI I(a,a)
  spicmv2.npz    SPICMV2 on a 500-row clustered block, 2 vectors
"""
import ctypes as C
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from grasp_amd import workloads as W  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "libgrasp_ref.so")


def ref_lib():
    # YZK keeps SAVEd powers of r per grid (yzk.f90:70-93): a fresh copy of
    # the library per grid
    tmp = tempfile.mkdtemp()
    path = os.path.join(tmp, "libgrasp_ref_copy.so")
    shutil.copy(REF, path)
    R = C.CDLL(path)
    R.ref_rinti.restype = C.c_double
    R.ref_yzk.restype = C.c_int
    R.ref_genintrk.restype = C.c_int
    return R


def ion(R, z, nmax, lmax, occ, amass):
    n, h = 590, 0.05
    rnt = 2e-6 / z
    parm = np.array(W.fermi_parameters(amass))
    orb = W.screened_orbitals(z, W.subshells(nmax, lmax), occupied=occ)
    R.ref_init(C.c_int(n), C.c_double(rnt), C.c_double(h), C.c_double(W.C_LIGHT), C.c_double(z), C.c_int(2),
               O.dp(parm))
    P = np.ascontiguousarray(orb.P)
    Q = np.ascontiguousarray(orb.Q)
    ints = [np.ascontiguousarray(a, dtype=np.int32) for a in (orb.mf, orb.nak, orb.nkl, orb.nkj)]
    R.ref_set_orbitals(C.c_int(P.shape[0]), C.c_int(P.shape[1]), O.dp(P), O.dp(Q), *[O.i32p(a) for a in ints])
    r, rp, rpor, zz = np.zeros(n + 10), np.zeros(n + 10), np.zeros(n + 10), np.zeros(n)
    R.ref_grid(O.dp(r), O.dp(rp), O.dp(rpor), O.dp(zz))
    d = dict(z=z, amass=amass, parm=parm, n=n, rnt=rnt, h=h, c=W.C_LIGHT, P=P, Q=Q, mf=ints[0], nak=ints[1],
             nkl=ints[2], nkj=ints[3], r=r, rp=rp, rpor=rpor, zz=zz)
    return d


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/libgrasp_ref.so not built: run oracle/build_ref.sh in a tree with the reference")
    # C III
    R = ref_lib()
    d = ion(R, 6.0, 4, 3, {1: 2, 2: 2}, 12.0)
    cap = 200000
    kst = np.zeros(25, dtype=np.int32)
    keys = np.zeros(cap, dtype=np.int32)
    vals = np.zeros(cap)
    nr = R.ref_genintrk(O.i32p(kst), O.i32p(keys), O.dp(vals), C.c_int(cap))
    j2max = int(np.max(d["nkj"]))
    nw = d["P"].shape[0]
    iab = np.zeros((nw, nw))
    for a in range(nw):
        for b in range(a, nw):
            if d["nak"][a] == d["nak"][b]:
                iab[a, b] = R.ref_rinti(C.c_int(a + 1), C.c_int(b + 1), C.c_int(0))
    yk = []
    ykey = []
    for (k, a, b) in [(0, 1, 1), (1, 1, 2), (2, 3, 5), (0, 4, 9)]:
        y = np.zeros(d["n"])
        mtp = R.ref_yzk(C.c_int(k), C.c_int(a), C.c_int(b), O.dp(y))
        yk.append(y)
        ykey.append((k, a, b, mtp))
    np.savez_compressed(os.path.join(HERE, "c3_fermi.npz"), **d, kstart=kst[: j2max + 2], keys=keys[:nr],
                        vals=vals[:nr], iab=iab, yzk=np.array(yk), yzk_key=np.array(ykey, dtype=np.int64))
    # Fe XVII
    R = ref_lib()
    d = ion(R, 26.0, 3, 2, {1: 2, 2: 8}, 56.0)
    nw = d["P"].shape[0]
    diag = np.array([R.ref_rinti(C.c_int(a + 1), C.c_int(a + 1), C.c_int(0)) for a in range(nw)])
    np.savez_compressed(os.path.join(HERE, "fe17_grid.npz"), **d, iaa=diag)
    # SPICMV2
    csr = W.clustered_block(500, 3, fill=0.1)
    B = np.asfortranarray(np.random.default_rng(0).standard_normal((500, 2)))
    Cr = np.zeros((500, 2), order="F")
    R.ref_spicmv2(C.c_int(500), C.c_int(2), O.i32p(csr.row_end.astype(np.int32)), O.i32p(csr.col),
                  O.dp(csr.val), C.c_int64(csr.nnz), O.dp(B), O.dp(Cr))
    np.savez_compressed(os.path.join(HERE, "spicmv2.npz"), row_end=csr.row_end, col=csr.col, val=csr.val, B=B,
                        C=Cr)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
