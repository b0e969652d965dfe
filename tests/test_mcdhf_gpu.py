"""The SCF program's seams (rmcdhf90) on the GPU vs the oracle.

rmcdhf90 is the second caller of the pieces rci90 uses: its MANEIG calls
GDVD(SPICMV2, ...) with its own rules (rmcdhf90/maneig.f90:78-200: LIM =
min(NCF, 2*NVECT+40), dense INIEST2 up to 4000 CSFs, warm starts from the
previous SCF iteration, levels ICCMIN), and its potentials call YZK pair by
pair (rmcdhf90/ypot.f90:75, xpot.f90:80).  Bars: Y^k functions bit-identical
to the oracle's YZK; eigenvalues within 1e-10 of the oracle relative to the
total level energy."""
import numpy as np
import pytest

import oracle as O
from grasp_amd import rci
from grasp_amd import workloads as W
from grasp_amd.rci import HamiltonianBlock

pytestmark = pytest.mark.gpu


def _eav(csr):
    return float(np.mean(csr.val[csr.row_end[1:] - 1]))


def _rel(got, want, eav):
    return np.abs(got - want) / np.maximum(np.abs(want + eav), 1.0)


def _shifted_vals(csr, eav):
    return np.where(np.isin(np.arange(csr.nnz), csr.row_end[1:] - 1), csr.val - eav, csr.val)


def test_yzk_batch_bit_identical(gpu):
    """Every Y^k(b, d) an SCF iteration over the Fe XVII n <= 5 orbitals can
    ask for (k up to 2 l_max, b <= d, triangle rule on l), including the
    K = 0 asymptote correction of YZK (b == d)."""
    z, n, h = 26.0, 590, 0.05
    rnt = 2e-6 / z
    orb = W.screened_orbitals(z, W.subshells(5, 4), occupied={1: 2, 2: 8})
    r, rp, rpor = rci.radgrd(n, rnt, h)
    g = O.Grid(n, rnt, h)
    o = O.Orbitals(orb.P, orb.Q, orb.mf, orb.nak, orb.nkl, orb.nkj)
    nw = len(orb.mf)
    ks, bs, ds = [], [], []
    for b in range(nw):
        for d in range(b, nw):
            lb, ld = int(orb.nkl[b]), int(orb.nkl[d])
            for k in range(abs(lb - ld), lb + ld + 1, 2):
                ks.append(k)
                bs.append(b + 1)
                ds.append(d + 1)
    yk, mtp = rci.yzk(r, rp, rpor, n, h, orb, ks, bs, ds)
    assert len(ks) > 300
    for q in range(len(ks)):
        want, wmtp = O.yzk(g, o, ks[q], bs[q] - 1, ds[q] - 1)
        assert mtp[q] == wmtp, (q, ks[q], bs[q], ds[q])
        assert np.array_equal(yk[q], want), (q, ks[q], bs[q], ds[q], np.abs(yk[q] - want).max())


@pytest.mark.parametrize("n,nvect,levels,seed", [(2500, 4, [1, 2, 4], 31), (700, 3, [3], 32)])
def test_mcdhf_maneig_dense(gpu, n, nvect, levels, seed):
    """NCF <= 4000: INIEST2 over the whole block is the answer (where rci90
    would already run DVDSON above 2000); levels may skip (ICCMIN = 1, 2, 4)."""
    csr = W.clustered_block(n, seed, fill=0.05, sorted_diag=True)
    eav = _eav(csr)
    h = HamiltonianBlock.from_csr(csr, shift=-eav)
    ev, evec, st = rci.maneig_mcdhf(h, nvect, levels)
    want, wvec, nl, nm, rc = O.maneig_mcdhf(csr.row_end, csr.col, csr.val, eav, nvect, levels, nthreads=8)
    assert rc == 0 and st["nloops"] == 0
    assert np.all(_rel(ev, want, eav) < 1e-10), (ev, want)
    assert np.max(np.abs(evec - wvec)) < 1e-6


def test_mcdhf_maneig_cold_and_warm_davidson(gpu):
    """NCF > 4000: cold start (INIEST2 over the leading 4000 CSFs, then GDVD
    with LIM = 2*NVECT+40), then a warm start of the next SCF iteration from
    the previous vectors of a slightly changed block; level 2 is left out of
    the warm start (its start column is the unit vector the header states)."""
    n, nvect, levels = 4600, 3, [1, 3]
    csr = W.clustered_block(n, 33, fill=0.04, sorted_diag=True)
    eav = _eav(csr)
    h = HamiltonianBlock.from_csr(csr, shift=-eav)
    ev, evec, st = rci.maneig_mcdhf(h, nvect, levels)
    want, wvec, nl, nm, rc = O.maneig_mcdhf(csr.row_end, csr.col, csr.val, eav, nvect, levels, nthreads=8)
    assert rc == 0 and st["ierr"] == 0 and st["nloops"] > 0
    assert np.all(_rel(ev, want, eav) < 1e-10), (ev, want)
    # the next SCF iteration: off-diagonals scaled by 1.01 (new orbitals)
    csr2 = W.clustered_block(n, 33, fill=0.04, sorted_diag=True)
    off = ~np.isin(np.arange(csr2.nnz), csr2.row_end[1:] - 1)
    csr2.val[off] *= 1.01
    eav2 = _eav(csr2)
    h2 = HamiltonianBlock.from_csr(csr2, shift=-eav2)
    ev2, evec2, st2 = rci.maneig_mcdhf(h2, nvect, levels, warm=(ev, evec))
    want2, wvec2, nl2, nm2, rc2 = O.maneig_mcdhf(csr2.row_end, csr2.col, csr2.val, eav2, nvect, levels,
                                                 warm=(want, wvec), nthreads=8)
    assert rc2 == 0 and st2["ierr"] == 0
    assert np.all(_rel(ev2, want2, eav2) < 1e-10), (ev2, want2)
    # residuals with the oracle operator
    A = O.spicmv2(csr2.row_end, csr2.col, _shifted_vals(csr2, eav2), evec2)
    assert np.max(np.linalg.norm(A - evec2 * ev2, axis=0)) < 1e-6
    h.free()
    h2.free()


def test_mcdhf_maneig_rejects_bad_levels(gpu):
    csr = W.clustered_block(300, 34, fill=0.05)
    h = HamiltonianBlock.from_csr(csr, shift=-_eav(csr))
    with pytest.raises(Exception, match="level"):
        rci.maneig_mcdhf(h, 2, [3])
