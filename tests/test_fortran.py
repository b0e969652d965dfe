"""Fortran host layer (grasp_amd/fortran): ISO_C_BINDING interfaces for the
whole C ABI, and an rci90-shaped host program (GCI_LOAD_HMAT, GCI_SPICMV2,
GCI_MANEIG_SOLVE, GCI_GENINTRK) checked against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from grasp_amd import build as B
from grasp_amd import rci
from grasp_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVE = os.path.join(B.FORT_BUILD, "gci_drive")
needs_fortran = pytest.mark.skipif(not os.path.exists(DRIVE), reason="Fortran host layer not built (no amdflang)")


def _c_functions():
    text = open(os.path.join(ROOT, "include", "graspci.h")).read()
    return set(re.findall(r"^\s*(?:int\s+|const char \*)(gci_\w+)\(", text, re.M))


def test_fortran_binds_every_c_entry_point():
    text = open(os.path.join(B.FORT_DIR, "gci_bind.f90")).read()
    bound = set(re.findall(r'bind\(C, name="(gci_\w+)"\)', text))
    assert _c_functions() <= bound, sorted(_c_functions() - bound)


@needs_fortran
def test_fortran_program_links_the_engine():
    out = subprocess.run(["ldd", DRIVE], capture_output=True, text=True).stdout
    lib = [ln for ln in out.splitlines() if "libgraspci.so" in ln]
    assert lib and "not found" not in lib[0]
    assert os.path.realpath(lib[0].split("=>")[1].split("(")[0].strip()) == os.path.realpath(B.OUT)


def _write_input(path, csr, eav, nvec, m, orb=None, grid=None):
    with open(path, "w") as f:
        f.write(f"{csr.n} {nvec} {m}\n{eav!r}\n")
        np.savetxt(f, csr.row_end[None, :], fmt="%d")
        np.savetxt(f, csr.col[None, :], fmt="%d")
        np.savetxt(f, csr.val[None, :], fmt="%.17e")
        if orb is None:
            f.write("0 0 0 0.05\n")
            return
        n, h, r, rp, rpor = grid
        nw, ld = orb.P.shape
        f.write(f"{n} {nw} {ld} {h!r}\n")
        for a in (r, rp, rpor, orb.P.reshape(-1), orb.Q.reshape(-1)):
            np.savetxt(f, np.asarray(a, dtype=np.float64)[None, :], fmt="%.17e")
        for a in (orb.mf, orb.nak, orb.nkl, orb.nkj):
            np.savetxt(f, np.asarray(a)[None, :], fmt="%d")


@needs_fortran
@pytest.mark.gpu
@pytest.mark.parametrize("n,nvec,with_orbitals", [(700, 3, True), (2600, 2, False)])
def test_fortran_host_matches_oracle(gpu, tmp_path, n, nvec, with_orbitals):
    csr = W.clustered_block(n, 31 + n, fill=0.06, sorted_diag=True)
    eav = float(np.mean(csr.val[csr.row_end[1:] - 1]))
    m = 3
    orb = grid = None
    if with_orbitals:
        z, rnt, h, ng = 6.0, 2e-6 / 6.0, 0.05, 590
        orb = W.screened_orbitals(z, W.subshells(4, 3), occupied={1: 2, 2: 2})
        r, rp, rpor = rci.radgrd(ng, rnt, h)
        grid = (ng, h, r, rp, rpor)
    fin, fout = tmp_path / "in.txt", tmp_path / "out.txt"
    _write_input(fin, csr, eav, nvec, m, orb, grid)
    print(f"gci_drive n={n} nvec={nvec} orbitals={with_orbitals}", flush=True)
    res = subprocess.run([DRIVE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr[-2000:]
    lines = open(fout).read().split("\n")
    vals = [ln for ln in lines if ln.strip()]
    # C = (H - EAV) B, B(i,k) = sin(0.37 i + 1.3 k)
    i = np.arange(1, n + 1)[:, None]
    k = np.arange(1, m + 1)[None, :]
    Bm = np.asfortranarray(np.sin(0.37 * i + 1.3 * k))
    got_c = np.array([float(x) for x in vals[: n * m]]).reshape(m, n).T
    shifted = csr.val.copy()
    shifted[csr.row_end[1:] - 1] -= eav
    want_c = O.spicmv2(csr.row_end, csr.col, shifted, Bm)
    scale = O.spicmv2(csr.row_end, csr.col, np.abs(shifted), np.abs(Bm))
    assert np.all(np.abs(got_c - want_c) <= 1e-13 * scale + 1e-300)
    # MANEIG eigenvalues of H - EAV, relative to total energies
    ev = np.array([float(x) for x in vals[n * m: n * m + nvec]])
    want_ev, *_ = O.maneig(csr.row_end, csr.col, csr.val, eav, nvec, nthreads=8)
    assert np.all(np.abs(ev - want_ev) / np.maximum(np.abs(want_ev + eav), 1.0) < 1e-10)
    nloops, nmv, ierr = (int(x) for x in vals[n * m + nvec].split())
    assert ierr == 0
    if with_orbitals:
        p = n * m + nvec + 1
        nint, j2max = (int(x) for x in vals[p].split())
        kst = np.array([int(x) for x in vals[p + 1: p + 3 + j2max]])
        pairs = [ln.split() for ln in vals[p + 3 + j2max: p + 3 + j2max + nint]]
        keys = np.array([int(a) for a, _ in pairs])
        rk = np.array([float(b) for _, b in pairs])
        g = O.Grid(grid[0], 2e-6 / 6.0, grid[1])
        o = O.Orbitals(orb.P, orb.Q, orb.mf, orb.nak, orb.nkl, orb.nkj)
        okst, okeys, ovals, *_ = O.genintrk(g, o)
        assert np.array_equal(kst, okst) and np.array_equal(keys, okeys)
        assert np.all(np.abs(rk - ovals) <= 1e-13 * np.maximum(np.abs(ovals), 1e-3))
