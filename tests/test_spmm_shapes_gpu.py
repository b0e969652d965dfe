"""SPICMV2 parity across the SpMM launch shapes (panel height, workgroup
size, prefetch depth, slices and entries per work item), every launch width
and every operator mode; the autotuner keeps the operator exact."""
import numpy as np
import pytest

import oracle as O
from grasp_amd import _lib as L
from grasp_amd import rci
from grasp_amd import workloads as W
from grasp_amd.rci import HamiltonianBlock

pytestmark = pytest.mark.gpu

SHAPES = [
    dict(R=1024, block=512, depth=4, nslice=32, target=65536),
    dict(R=2048, block=256, depth=2, nslice=64, target=131072),
    dict(R=1024, block=1024, depth=2, nslice=1, target=4096),   # one slice and 16 steps per item
    dict(R=2048, block=512, depth=4, nslice=3, target=8192),
]


@pytest.fixture
def shape_guard(gpu):
    saved = rci.spmm_shape()
    yield
    rci.set_spmm_shape(**saved)


def _check(h, csr, m, mode, seed):
    rng = np.random.default_rng(seed)
    B = np.asfortranarray(rng.standard_normal((csr.n, m)))
    got = h.spicmv2(B, mode)
    want = O.spicmv2(csr.row_end, csr.col, csr.val, B)
    scale = O.spicmv2(csr.row_end, csr.col, np.abs(csr.val), np.abs(B))
    # FP64 and hi+lo: backward-error bound 1e-13 (unordered atomics);
    # bulk: fp32 values and fp32-rounded B
    tol = 2.5e-7 if mode == L.GCI_OP_SPLIT_BULK else 1e-13
    err = np.abs(got - want)
    assert np.all(err <= tol * scale + 1e-300), (m, mode, float(np.max(err / (scale + 1e-300))))


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"R{s['R']}-b{s['block']}-d{s['depth']}-s{s['nslice']}")
def test_spmm_shapes_all_widths(shape_guard, shape):
    rci.set_spmm_shape(**shape)
    csr = W.clustered_block(5200, 31, fill=0.04)
    h32 = HamiltonianBlock.from_csr(csr, precision=L.GCI_SPLIT32)
    h64 = HamiltonianBlock.from_csr(csr)
    for m in (1, 2, 3, 4, 5, 8, 9):
        _check(h64, csr, m, L.GCI_OP_FP64, 100 + m)
        _check(h32, csr, m, L.GCI_OP_SPLIT_FULL, 200 + m)
        _check(h32, csr, m, L.GCI_OP_SPLIT_BULK, 300 + m)


@pytest.mark.parametrize("fill,n", [(0.9, 1500), (0.0004, 9000)], ids=["dense", "single-entry-segments"])
def test_spmm_segment_extremes(shape_guard, fill, n):
    # a nearly dense block (segments of whole panel height, one slice per
    # panel column group) and a very sparse one (segments of one entry,
    # slices that are one step long; in the full layouts of the split
    # matrix, row gaps far above 3 that take escape entries)
    rci.set_spmm_shape(R=1024, block=512, depth=4, nslice=8, target=16384)
    csr = W.clustered_block(n, 41, fill=fill)
    h = HamiltonianBlock.from_csr(csr)
    h32 = HamiltonianBlock.from_csr(csr, precision=L.GCI_SPLIT32)
    for m in (1, 4, 8):
        _check(h, csr, m, L.GCI_OP_FP64, 400 + m)
        _check(h32, csr, m, L.GCI_OP_SPLIT_FULL, 500 + m)
        _check(h32, csr, m, L.GCI_OP_SPLIT_BULK, 600 + m)
        _check(h32, csr, m, L.GCI_OP_FP64, 700 + m)


def _lower_csr(n, entries, seed):
    """LowerCSR with the given strictly-lower (row, column) pairs (0-based)."""
    rng = np.random.default_rng(seed)
    by_row = {}
    for i, j in entries:
        assert i > j
        by_row.setdefault(i, []).append(j)
    row_end, col, val = [0], [], []
    for i in range(n):
        for j in sorted(by_row.get(i, [])):
            col.append(j + 1)
            val.append(rng.normal(0.0, 1e-2))
        col.append(i + 1)
        val.append(rng.normal(-338.0, 5.0))
        row_end.append(len(col))
    return W.LowerCSR(n, np.array(row_end, np.int64), np.array(col, np.int32), np.array(val))


def test_spmm_full_layout_escapes(shape_guard):
    """Row gaps at and around the two-bit code limit in a full layout: 1, 2
    and 3 (codes, no escape), 4 and 5 (one escape entry holding the row),
    255, 256 and 511 (one escape each), a segment whose first row is the
    first of its panel, one far into its panel, and one in a later panel.
    The running row must land on every entry exactly, in all operator modes
    of the split matrix and in the FP64 matrix's lower layout."""
    rci.set_spmm_shape(R=2048, block=512, depth=8, nslice=16, target=131072)
    n = 5000
    entries = [(r, 5) for r in (6, 7, 9, 12, 16, 21, 276, 532, 1043, 1538, 2047)]
    entries += [(r, 0) for r in (300, 600, 1000, 2000, 2047)]
    entries += [(2040, 1000), (4100, 3000), (4999, 3000), (4999, 4998), (2048, 2047)]
    csr = _lower_csr(n, entries, 77)
    h32 = HamiltonianBlock.from_csr(csr, precision=L.GCI_SPLIT32)
    h64 = HamiltonianBlock.from_csr(csr)
    for m in (1, 4, 5):
        _check(h32, csr, m, L.GCI_OP_SPLIT_BULK, 800 + m)
        _check(h32, csr, m, L.GCI_OP_SPLIT_FULL, 810 + m)
        _check(h32, csr, m, L.GCI_OP_FP64, 820 + m)
        _check(h64, csr, m, L.GCI_OP_FP64, 830 + m)


def test_spmm_autotune_keeps_parity(shape_guard):
    csrs = [W.clustered_block(n, 50 + n, fill=0.05) for n in (1800, 4100)]
    blocks = [HamiltonianBlock.from_csr(c, precision=L.GCI_SPLIT32) for c in csrs]
    shape, ms = rci.autotune(blocks, 4, L.GCI_OP_SPLIT_BULK, reps=2)
    assert ms > 0 and rci.spmm_shape() == shape
    for c, h in zip(csrs, blocks):
        _check(h, c, 4, L.GCI_OP_SPLIT_BULK, 7)
        _check(h, c, 4, L.GCI_OP_SPLIT_FULL, 8)


def test_spmm_layout_overhead(shape_guard):
    # a full layout pads each slice to its longest segment and adds an escape
    # entry for every row gap above 3; on a clustered block the stream stays
    # close to the laid-out entries
    rci.set_spmm_shape(R=1024)
    csr = W.clustered_block(6000, 61, fill=0.05)
    h = HamiltonianBlock.from_csr(csr, precision=L.GCI_SPLIT32)
    lay = rci.spmm_layout(h)
    assert lay["slices"] >= lay["segments"] / 64
    assert lay["stored"] % 256 == 0
    assert 1.0 <= lay["overhead"] < 1.5, lay
