"""The no-allocation SpMM entry point (gci_spmm_device_noalloc) against the
allocating one (gci_spmm_device) through the C ABI: two calls in a row, the
second on a smaller input, in every operator mode; and its size limit.

Both run the same kernels; the SpMM's global atomics sum in a run-dependent
order, so results agree to the SpMM's parity bar, 1e-13 of the
backward-error bound sum_j |H_ij| |x_j| (the oracle's SPICMV2 on |H|, |x|)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from grasp_amd import _lib as L
from grasp_amd import rci
from grasp_amd import workloads as W

pytestmark = pytest.mark.gpu


_HIP = None


def _hip():
    """The HIP runtime libgraspci is bound to (loaded with it), for the
    caller's own device buffers."""
    global _HIP
    if _HIP is None:
        L.lib()
        _HIP = C.CDLL("libamdhip64.so.7")
    return _HIP


class _Dev:
    """A caller-owned device copy of a float64 array."""

    def __init__(self, a):
        self.shape = a.shape
        self.nbytes = a.size * 8
        self.p = C.c_void_p()
        assert _hip().hipMalloc(C.byref(self.p), C.c_size_t(self.nbytes)) == 0
        host = np.ascontiguousarray(a, dtype=np.float64)
        assert _hip().hipMemcpy(self.p, host.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1) == 0

    def data_ptr(self):
        return self.p.value

    def get(self):
        out = np.empty(self.shape)
        assert _hip().hipDeviceSynchronize() == 0
        assert _hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(self.nbytes), 2) == 0
        return out

    def __del__(self):
        if self.p:
            _hip().hipFree(self.p)
            self.p = C.c_void_p()


def _dev(a):
    return _Dev(a)


def _apply(fn, h, Bd, Cd, m, mode):
    assert _hip().hipDeviceSynchronize() == 0
    L.check(fn(h.handle, C.c_int(m), C.c_void_p(Bd.data_ptr()), C.c_int(h.n), C.c_void_p(Cd.data_ptr()),
               C.c_int(h.n), C.c_int(mode)))
    return Cd.get().T  # (m, n) row-major = n x m column-major (get() waits for the handle's stream)


def _bound(csr, eav, B):
    diag = np.zeros(csr.nnz, dtype=bool)
    diag[csr.row_end[1:] - 1] = True
    held = np.where(diag, csr.val - eav, csr.val)
    return O.spicmv2(csr.row_end, csr.col, np.abs(held), np.abs(B)), held


@pytest.mark.parametrize("precision,modes", [(L.GCI_SPLIT32, (L.GCI_OP_SPLIT_BULK, L.GCI_OP_SPLIT_FULL,
                                                              L.GCI_OP_FP64)),
                                             (L.GCI_FP64, (L.GCI_OP_FP64,))])
def test_noalloc_matches_allocating_entry_point(gpu, precision, modes):
    lib = L.lib()
    rng = np.random.default_rng(11)
    for n, m, seed in ((5000, 5, 31), (1200, 2, 32)):  # the second call is smaller
        csr = W.clustered_block(n, seed, fill=0.04)
        eav = float(np.mean(csr.val[csr.row_end[1:] - 1]))
        h = rci.HamiltonianBlock.from_csr(csr, shift=-eav, precision=precision)
        rci.spmm_layout(h)  # the layout and its device items, prepared beforehand
        B = np.asfortranarray(rng.standard_normal((n, m)))
        bound, held = _bound(csr, eav, B)
        Bd = _dev(B.T)
        for mode in modes:
            want = _apply(lib.gci_spmm_device, h, Bd, _dev(np.zeros((m, n))), m, mode)
            got = _apply(lib.gci_spmm_device_noalloc, h, Bd, _dev(np.full((m, n), np.nan)), m, mode)
            assert np.all(np.isfinite(got)), (n, mode)
            assert np.all(np.abs(got - want) <= 1e-13 * bound + 1e-300), (n, mode)
            if mode != L.GCI_OP_SPLIT_BULK:
                ref = O.spicmv2(csr.row_end, csr.col, held, B)
                assert np.all(np.abs(got - ref) <= 1e-13 * bound + 1e-300), (n, mode)
        h.free()


def test_noalloc_refuses_oversize_and_unprepared(gpu):
    lib = L.lib()
    # 600000 rows x 8 columns of fp32 x and fp64 y staging exceed the 48 MiB
    # (a diagonal block with one off-diagonal entry)
    n = 600000
    iendc = np.concatenate([[0], np.arange(1, n + 1) + (np.arange(n) >= 1)]).astype(np.int64)
    irow = np.concatenate([[1, 1, 2], np.arange(3, n + 1)]).astype(np.int32)
    emt = np.concatenate([[-1.0, 0.01, -2.0], -np.arange(3, n + 1, dtype=np.float64)])
    h = rci.HamiltonianBlock.from_rows(iendc, irow, emt, precision=L.GCI_SPLIT32)
    rci.spmm_layout(h)
    x = _dev(np.zeros((8, n)))
    y = _dev(np.zeros((8, n)))
    code = lib.gci_spmm_device_noalloc(h.handle, C.c_int(8), C.c_void_p(x.data_ptr()), C.c_int(n),
                                       C.c_void_p(y.data_ptr()), C.c_int(n), C.c_int(L.GCI_OP_SPLIT_BULK))
    assert code == L.GCI_ERR_ARG and b"scratch" in lib.gci_last_error()
    # within the limit on the same matrix: two columns
    B = np.random.default_rng(3).standard_normal((n, 2))
    got = _apply(lib.gci_spmm_device_noalloc, h, _dev(B.T), _dev(np.zeros((2, n))), 2, L.GCI_OP_SPLIT_FULL)
    want = _apply(lib.gci_spmm_device, h, _dev(B.T), _dev(np.zeros((2, n))), 2, L.GCI_OP_SPLIT_FULL)
    assert np.all(np.abs(got - want) <= 1e-13 * np.abs(want) + 1e-300)
    h.free()
    # a matrix whose layout was never built is refused, not laid out
    csr = W.clustered_block(700, 33, fill=0.05)
    h2 = rci.HamiltonianBlock.from_csr(csr)
    x = _dev(np.zeros((1, csr.n)))
    y = _dev(np.zeros((1, csr.n)))
    code = lib.gci_spmm_device_noalloc(h2.handle, C.c_int(1), C.c_void_p(x.data_ptr()), C.c_int(csr.n),
                                       C.c_void_p(y.data_ptr()), C.c_int(csr.n), C.c_int(L.GCI_OP_FP64))
    assert code == L.GCI_ERR_ARG and b"gci_spmm_layout" in lib.gci_last_error()
    h2.free()
