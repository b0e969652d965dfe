"""SPICMV2 parity: HIP symmetric SpMM vs the oracle (lib9290/spicmv2.f90)."""
import numpy as np
import pytest

import oracle as O
from grasp_amd import _lib as L
from grasp_amd import workloads as W
from grasp_amd.rci import HamiltonianBlock

pytestmark = pytest.mark.gpu


def _bound(csr, B):
    """Per-element backward-error scale sum_j |H_ij| |B_j| (|H| symmetric)."""
    A = np.abs(csr.dense()) if csr.n <= 3000 else None
    if A is not None:
        return A @ np.abs(B)
    return O.spicmv2(csr.row_end, csr.col, np.abs(csr.val), np.abs(B))


@pytest.mark.parametrize("n,fill,seed,m", [(1, 0.0, 1, 1), (2, 0.5, 2, 1), (97, 0.2, 3, 1),
                                           (640, 0.05, 4, 3), (2500, 0.08, 5, 8), (4000, 0.02, 6, 5)])
def test_spmm_fp64_matches_oracle(gpu, n, fill, seed, m):
    csr = W.clustered_block(n, seed, fill=fill)
    h = HamiltonianBlock.from_csr(csr)
    rng = np.random.default_rng(seed)
    B = np.asfortranarray(rng.standard_normal((n, m)))
    got = h.spicmv2(B)
    want = O.spicmv2(csr.row_end, csr.col, csr.val, B)
    scale = _bound(csr, B)
    # FP64 with unordered atomic accumulation: backward-error bound 1e-13
    assert np.all(np.abs(got - want) <= 1e-13 * scale + 1e-300)


def test_spmm_roundtrip_download(gpu):
    csr = W.clustered_block(300, 11, fill=0.1)
    h = HamiltonianBlock.from_csr(csr, shift=-2.5)
    iendc, irow, emt = h.download()
    assert np.array_equal(iendc, csr.row_end)
    assert np.array_equal(irow, csr.col)
    d = csr.row_end[1:] - 1
    want = csr.val.copy()
    want[d] += -2.5
    assert np.array_equal(emt, want)


@pytest.mark.parametrize("mode", [L.GCI_OP_SPLIT_BULK, L.GCI_OP_SPLIT_FULL])
def test_spmm_split_modes(gpu, mode):
    csr = W.clustered_block(3000, 21, fill=0.05)
    h = HamiltonianBlock.from_csr(csr, precision=L.GCI_SPLIT32)
    rng = np.random.default_rng(3)
    B = np.asfortranarray(rng.standard_normal((csr.n, 8)))
    got = h.spicmv2(B, mode)
    want = O.spicmv2(csr.row_end, csr.col, csr.val, B)
    scale = _bound(csr, B)
    # bulk: fp32 off-diagonals and fp32-rounded B (2^-23 each), FP64 diagonal
    # and accumulation; full: hi+lo carries 48 bits of each value
    tol = 2.5e-7 if mode == L.GCI_OP_SPLIT_BULK else 1e-13
    assert np.all(np.abs(got - want) <= tol * scale)


def test_spmm_rejects_bad_rows(gpu):
    # diagonal not last
    with pytest.raises(L.GciError):
        HamiltonianBlock.from_rows(np.array([0, 2]), np.array([1, 1]), np.array([1.0, 2.0]))
    # column above the diagonal
    with pytest.raises(L.GciError):
        HamiltonianBlock.from_rows(np.array([0, 1, 3]), np.array([1, 3, 2]), np.array([1.0, 2.0, 3.0]))
