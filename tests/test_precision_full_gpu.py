"""The precision form at full size: one 3e5-row block at 5% lower fill
(2.25e9 stored lower entries), split FP32 storage, 8 eigenpairs from the
8-vector block Davidson with FP64 refinement.

The oracle cannot finish an eigensolve of this size in a test, so the
results are checked through properties that do not depend on size: each
eigenpair's residual with the FP64 operator (the oracle's SPICMV2 on the
block's FP64 values, threaded row-cyclically as rci_mpi splits it,
lib9290/spicmv2.f90, mpi90/spicmvmpi.f90:61-75) is at most tol |E|, and the
eigenvectors are orthonormal."""
import os
import sys
import time

import numpy as np
import pytest

import oracle as O
from grasp_amd import _lib as L
from grasp_amd import rci, synth

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.mark.timeout(420)
def test_precision_form_full_size_properties(gpu, record_property):
    n, fill, seed, nev, tol = 300000, 0.05, 2024, 8, 1e-9
    t0 = time.time()
    h, eav = synth.synth_block(n, seed, fill=fill, precision=L.GCI_SPLIT32)
    _log(f"precision full: generated n={h.n} nnz_L={h.nnz_offdiag} in {time.time() - t0:.1f}s")
    try:
        assert h.nnz_offdiag > 2.2e9
        t0 = time.time()
        ev, V, st = rci.batch_solve([h], [nev], tol, [eav])
        ev, V = ev[0], np.asfortranarray(V[0])
        _log(f"precision full: solved in {time.time() - t0:.1f}s, stats {st[0]}")
        assert st[0]["ierr"] == 0 and st[0]["refine_loops"] >= 1
        assert np.all(np.diff(ev) >= 0.0)
        assert np.max(np.abs(V.T @ V - np.eye(nev))) < 1e-10
        t0 = time.time()
        iendc, irow, emt = h.download()
        _log(f"precision full: downloaded in {time.time() - t0:.1f}s")
    finally:
        h.free()
    t0 = time.time()
    threads = min(16, os.cpu_count() or 1)
    HV = O.spicmv2(iendc, irow, emt, V, nthreads=threads)
    del iendc, irow, emt
    r = np.linalg.norm(HV - V * ev, axis=0)
    rel = r / np.abs(ev + eav)
    _log(f"precision full: FP64 residuals in {time.time() - t0:.1f}s on {threads} threads: {rel}")
    record_property("max_rel_residual", float(rel.max()))
    record_property("eigenvalues_total", [float(x) for x in ev + eav])
    assert np.all(rel <= tol), rel
