"""Davidson / MANEIG parity: device GDVD vs the oracle (libdvd90, rci90/maneig.f90)."""
import os

import numpy as np
import pytest

import oracle as O
from grasp_amd import _lib as L
from grasp_amd import workloads as W
from grasp_amd.rci import HamiltonianBlock, gdvd, maneig

pytestmark = pytest.mark.gpu


def _eav(csr):
    return float(np.mean(csr.val[csr.row_end[1:] - 1]))


def _rel(got, want, eav):
    # relative to the total level energy E = EVAL + EAV (rci90/matrix.f90:198-204)
    return np.abs(got - want) / np.maximum(np.abs(want + eav), 1.0)


@pytest.mark.parametrize("n,nvec,seed", [(40, 2, 1), (700, 4, 2), (1999, 3, 3)])
def test_maneig_dense_branch(gpu, n, nvec, seed):
    csr = W.clustered_block(n, seed, fill=0.08)
    eav = _eav(csr)
    h = HamiltonianBlock.from_csr(csr, shift=-eav)
    ev, evec, st = maneig(h, nvec)
    want, wvec, *_ = O.maneig(csr.row_end, csr.col, csr.val, eav, nvec)
    assert np.all(_rel(ev, want, eav) < 1e-10)
    # eigenvectors: same sign convention, so compare directly
    assert np.max(np.abs(evec - wvec)) < 1e-6


@pytest.mark.parametrize("n,nvec,seed", [(2600, 2, 4), (5000, 4, 5)])
def test_maneig_davidson_branch(gpu, n, nvec, seed):
    csr = W.clustered_block(n, seed, fill=0.05, sorted_diag=True)
    eav = _eav(csr)
    h = HamiltonianBlock.from_csr(csr, shift=-eav)
    ev, evec, st = maneig(h, nvec)
    want, wvec, nl, nm, rc = O.maneig(csr.row_end, csr.col, csr.val, eav, nvec, nthreads=8)
    assert rc == 0 and st["ierr"] == 0
    assert np.all(_rel(ev, want, eav) < 1e-10)
    # residuals of the returned vectors with the oracle operator
    A = O.spicmv2(csr.row_end, csr.col, np.where(np.isin(np.arange(csr.nnz), csr.row_end[1:] - 1),
                                                   csr.val - eav, csr.val), evec)
    r = A - evec * ev
    # TSTSEL may stop on CRITC/CRITE before CRITR (dvdson.f90:1328-1370)
    assert np.max(np.linalg.norm(r, axis=0)) < 1e-6


def test_gdvd_block_matches_eigh(gpu):
    n, nvec = 1500, 4
    csr = W.clustered_block(n, 9, fill=0.1)
    eav = _eav(csr)
    h = HamiltonianBlock.from_csr(csr, shift=-eav)
    ev, evec, res, st = gdvd(h, nvec, mblock=nvec, init=None)
    dense = csr.dense() - eav * np.eye(n)
    w = np.linalg.eigh(dense)[0][:nvec]
    assert np.all(_rel(ev, w, eav) < 1e-10)
    # TSTSEL may stop on CRITC/CRITE before CRITR (dvdson.f90:1328-1370)
    assert np.all(res < 1e-6)
    V = evec
    assert np.max(np.abs(V.T @ V - np.eye(nvec))) < 1e-10


def test_gdvd_split_with_refinement(gpu):
    n, nvec = 3000, 4
    csr = W.clustered_block(n, 10, fill=0.05, sorted_diag=True)
    eav = _eav(csr)
    h = HamiltonianBlock.from_csr(csr, shift=-eav, precision=L.GCI_SPLIT32)
    ev, evec, res, st = gdvd(h, nvec, mblock=nvec, mode=L.GCI_OP_SPLIT_BULK, refine=True)
    want, *_ = O.maneig(csr.row_end, csr.col, csr.val, eav, nvec, nthreads=8)
    assert np.all(_rel(ev, want, eav) < 1e-10)
    assert st["refine_loops"] >= 1


def test_maneig_new_column_below_rounding_unit(gpu):
    # A block whose INIEST2 start (eigenvectors of the leading 2000 rows)
    # gets, at the second iteration, a new basis column that couples to the
    # wanted Ritz vectors only at 1e-10: the lowest Ritz values do not move
    # by one rounding unit, which must not count as CRITE = 1e-17 being met
    # (the residuals are still 1e-4 there).
    from grasp_amd import synth
    n, fill, seed = synth.batched_workload(64, 2024)[55]
    h, eav = synth.synth_block(n, seed, fill=fill, precision=L.GCI_FP64)
    try:
        ev, evec, st = maneig(h, 4)
        r = h.spicmv2(np.asfortranarray(evec)) - evec * ev
        rel = np.linalg.norm(r, axis=0) / np.abs(ev + eav)
        assert st["nloops"] > 2 and st["ierr"] == 0
        assert np.all(rel < 1e-8), rel
    finally:
        h.free()


def _held(csr, eav):
    """Values of the block as MANEIG holds it (diagonal shifted by -EAV)."""
    diag = np.zeros(csr.nnz, dtype=bool)
    diag[csr.row_end[1:] - 1] = True
    return np.where(diag, csr.val - eav, csr.val)


def _fp64_residuals(csr, eav, V, ev):
    """||H v - lambda v|| per column with the oracle's FP64 SPICMV2."""
    A = O.spicmv2(csr.row_end, csr.col, _held(csr, eav), V, nthreads=8)
    return np.linalg.norm(A - V * ev, axis=0)


@pytest.mark.parametrize("nev", [12, 16])
def test_maneig_nvec_above_ten(gpu, nev):
    # LIM = MIN(NCF, 2*NVECMX+60) = 84 for 12 eigenpairs and 92 for 16, the
    # largest NVECMX the engine takes (maneig.f90:389); with 16, INITDVD
    # builds NUME + 1 = 17 columns (S = B^T AB in two blocks of the Gram
    # kernel's 16-column tile) and the Rayleigh-Ritz step computes 17 Ritz
    # vectors
    csr = W.clustered_block(3000, 12, fill=0.05, sorted_diag=True)
    eav = _eav(csr)
    h = HamiltonianBlock.from_csr(csr, shift=-eav)
    ev, evec, st = maneig(h, nev)
    want, wvec, nl, nm, rc = O.maneig(csr.row_end, csr.col, csr.val, eav, nev, nthreads=8)
    assert rc == 0 and st["ierr"] == 0
    assert np.all(_rel(ev, want, eav) < 1e-10), _rel(ev, want, eav)
    # DVDSON stops a vector on any of CRITE, CRITC, CRITR, so a residual can
    # stay above CRITR: the same as the oracle's (one vector at 4.2e-6 for
    # 16 pairs on this block), else below 1e-6
    ours, theirs = _fp64_residuals(csr, eav, evec, ev), _fp64_residuals(csr, eav, wvec, want)
    assert np.all(ours < np.maximum(1e-6, 1.01 * theirs)), (ours, theirs)


def test_davidson_rejects_nvec_beyond_limit(gpu):
    csr = W.clustered_block(3000, 13, fill=0.02)
    h = HamiltonianBlock.from_csr(csr, shift=-_eav(csr))
    with pytest.raises(L.GciError) as e:
        maneig(h, 17)
    assert e.value.code == L.GCI_ERR_ARG and "1..16" in str(e.value)


@pytest.mark.parametrize("mblock", [1, 4])
def test_gdvd_small_lim_restarts(gpu, mblock, record_property):
    """LIM = NVEC + 6 forces OVFLOW restarts (dvdson.f90:743-750,
    1179-1256): the basis is rotated onto the REST lowest Ritz vectors each
    time KPASS reaches LIM.  Same start vectors and LIM as the oracle."""
    n, nvec, lim = 3000, 4, 10
    csr = W.clustered_block(n, 14, fill=0.05, sorted_diag=True)
    eav = _eav(csr)
    held = _held(csr, eav)
    init = O.iniest2(csr.row_end, csr.col, held, nvec)
    diag = held[csr.row_end[1:] - 1]
    want, wvec, onl, onm, oerr, ores = O.gdvd(csr.row_end, csr.col, held, diag, nvec, lim, init,
                                               mblock=mblock, nthreads=8)
    h = HamiltonianBlock.from_csr(csr, shift=-eav)
    ev, evec, res, st = gdvd(h, nvec, lim=lim, mblock=mblock, init=init)
    _, _, _, st_free = gdvd(h, nvec, mblock=mblock, init=init)  # LIM = 68: no restart
    assert oerr == 0 and st["ierr"] == 0
    assert np.all(_rel(ev, want, eav) < 1e-10), _rel(ev, want, eav)
    # KPASS grows by at most MBLOCK per iteration from NIV = NVEC, so more
    # than (LIM - NVEC) / MBLOCK iterations means at least one restart
    assert st["nloops"] > (lim - nvec) // mblock and onl > (lim - nvec) // mblock
    if mblock > 1:
        # the 4-vector block fills LIM = 10 in two iterations; the oracle
        # needs 19 iterations against 7 with LIM = 68 on this block
        assert st["nloops"] > st_free["nloops"]
    record_property("nloops", st["nloops"])
    record_property("oracle_nloops", onl)
    assert np.max(_fp64_residuals(csr, eav, evec, ev)) < 1e-6
    assert np.max(np.abs(evec.T @ evec - np.eye(nvec))) < 1e-10


def test_gdvd_maxiter_reports_noconv(gpu):
    """MAXITER reached: DVDSON's IERR 2048 (dvdson.f90:817) becomes
    GCI_ERR_NOCONV."""
    csr = W.clustered_block(2600, 15, fill=0.05, sorted_diag=True)
    h = HamiltonianBlock.from_csr(csr, shift=-_eav(csr))
    with pytest.raises(L.GciError) as e:
        gdvd(h, 3, maxiter=1)
    assert e.value.code == L.GCI_ERR_NOCONV


def _golden_block(case):
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_blocks.npz"))
    spec = {"config2_2e4": (20000, 21, 0.05), "config3_3e4": (30000, 31, 0.05)}[case]
    csr = W.clustered_block(spec[0], spec[1], fill=spec[2], sorted_diag=True)
    fp = np.array([csr.n, csr.nnz, float(np.sum(csr.val)), float(np.sum(csr.val * csr.val))])
    assert np.array_equal(fp[:2], g[f"{case}_fingerprint"][:2])
    assert np.allclose(fp[2:], g[f"{case}_fingerprint"][2:], rtol=1e-12, atol=0.0)
    eav = float(g[f"{case}_eav"][0])
    assert eav == _eav(csr)
    return csr, eav, g[f"{case}_eval"]


def test_maneig_plain_fp64_config2_block(gpu, record_property):
    """The plain FP64 drop-in on a 2e4-row block (the Fe XVII n <= 6 block
    size): INIEST2 + GDVD(SPICMV2) with MANEIG's criteria, against the
    oracle's MANEIG eigenvalues (tests/golden/oracle_blocks.npz)."""
    csr, eav, want = _golden_block("config2_2e4")
    h = HamiltonianBlock.from_csr(csr, shift=-eav)
    ev, evec, st = maneig(h, 4)
    rel = _rel(ev, want, eav)
    record_property("max_rel_err", float(rel.max()))
    record_property("nloops", st["nloops"])
    assert st["ierr"] == 0
    assert np.all(rel < 1e-10), rel
    assert np.max(_fp64_residuals(csr, eav, evec, ev)) < 1e-6


def test_precision_form_at_oracle_size(gpu, record_property):
    """The precision form's solve (split FP32 storage, 8-vector block
    Davidson over the bulk FP32 operator with FP64 accumulation, FP64
    hi+lo refinement) for 8 eigenpairs of a 3e4-row block, against the
    oracle's FP64 MANEIG eigenvalues at 1e-9 relative."""
    csr, eav, want = _golden_block("config3_3e4")
    h = HamiltonianBlock.from_csr(csr, shift=-eav, precision=L.GCI_SPLIT32)
    ev, evec, res, st = gdvd(h, 8, mblock=8, mode=L.GCI_OP_SPLIT_BULK, refine=True)
    rel = _rel(ev, want, eav)
    record_property("max_rel_err", float(rel.max()))
    record_property("nloops", st["nloops"])
    record_property("refine_loops", st["refine_loops"])
    assert st["ierr"] == 0 and st["refine_loops"] >= 1
    assert np.all(rel < 1e-9), rel
    r = _fp64_residuals(csr, eav, evec, ev)
    assert np.max(r) < 1e-6
    assert np.max(np.abs(evec.T @ evec - np.eye(8))) < 1e-10
