"""Batched form and the continuous-batching server vs the oracle.

Every eigenvalue must be within the requested relative tolerance of the
FP64 oracle (dense eigh of the same block), relative to the total level
energy E = EVAL + EAV."""
import time

import numpy as np
import pytest

import oracle as O
from grasp_amd import _lib as L
from grasp_amd import rci
from grasp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _blocks(sizes, seed, fills):
    out = []
    for i, (n, f) in enumerate(zip(sizes, fills)):
        csr = W.clustered_block(int(n), seed + i, fill=f)
        eav = float(np.mean(csr.val[csr.row_end[1:] - 1]))
        out.append((csr, eav))
    return out


def _reference(csr, eav, nev):
    w = np.linalg.eigvalsh(csr.dense() - eav * np.eye(csr.n))
    return w[:nev]


def test_batch_solve_mixed_sizes(gpu):
    sizes = [120, 350, 900, 1600, 2400, 240, 3000, 700]
    fills = [0.1, 0.08, 0.05, 0.04, 0.03, 0.1, 0.02, 0.06]
    nev = [4, 2, 4, 3, 4, 1, 4, 2]
    tol = 1e-9
    data = _blocks(sizes, 100, fills)
    blocks = [rci.HamiltonianBlock.from_csr(c, shift=-e, precision=L.GCI_SPLIT32) for c, e in data]
    evals, evecs, stats = rci.batch_solve(blocks, nev, tol, [e for _, e in data])
    for b, ((csr, eav), k) in enumerate(zip(data, nev)):
        want = _reference(csr, eav, k)
        rel = np.abs(evals[b] - want) / np.abs(want + eav)
        assert np.all(rel < tol), (b, rel)
        assert stats[b]["ierr"] == 0
        assert stats[b]["refine_loops"] >= 1
        V = evecs[b]
        assert np.max(np.abs(V.T @ V - np.eye(k))) < 1e-8


def test_batch_solve_fp64_blocks(gpu):
    data = _blocks([400, 1500], 200, [0.1, 0.05])
    blocks = [rci.HamiltonianBlock.from_csr(c, shift=-e) for c, e in data]
    evals, evecs, stats = rci.batch_solve(blocks, [3, 3], 1e-10, [e for _, e in data])
    for b, (csr, eav) in enumerate(data):
        want = _reference(csr, eav, 3)
        assert np.all(np.abs(evals[b] - want) / np.abs(want + eav) < 1e-10)


def test_server_alone_vs_mix(gpu):
    """Each request's result is the same alone and in the mix (within tol)."""
    tol = 1e-9
    sizes = [300, 1200, 2500, 800, 150, 1900]
    data = _blocks(sizes, 300, [0.08, 0.05, 0.03, 0.06, 0.1, 0.04])
    nev = [2, 4, 3, 4, 1, 2]
    # alone: one request at a time
    alone = []
    srv = rci.Server(8 << 30, max_active=8)
    for (csr, eav), k in zip(data, nev):
        h = rci.HamiltonianBlock.from_csr(csr, shift=-eav, precision=L.GCI_SPLIT32)
        rid = srv.submit(h, k, tol, eav)
        srv.drain()
        alone.append(srv.poll(rid)[0])
    # in the mix: all submitted together, staggered arrivals
    rids = []
    for (csr, eav), k in zip(data, nev):
        h = rci.HamiltonianBlock.from_csr(csr, shift=-eav, precision=L.GCI_SPLIT32)
        rids.append(srv.submit(h, k, tol, eav))
        time.sleep(0.002)
    srv.drain()
    mix = [srv.poll(r) for r in rids]
    srv.close()
    for i, ((csr, eav), k) in enumerate(zip(data, nev)):
        want = _reference(csr, eav, k)
        for got in (alone[i], mix[i][0]):
            assert np.all(np.abs(got - want) / np.abs(want + eav) < tol), i
        assert np.all(np.abs(alone[i] - mix[i][0]) / np.abs(want + eav) < tol)
        assert mix[i][2] > 0.0


def test_server_budget_rejects_oversize(gpu):
    csr = W.clustered_block(2000, 9, fill=0.05)
    srv = rci.Server(1 << 20, max_active=2)  # 1 MiB budget
    h = rci.HamiltonianBlock.from_csr(csr, precision=L.GCI_SPLIT32)
    with pytest.raises(L.GciError) as e:
        srv.submit(h, 2, 1e-8)
    assert e.value.code == -3
    srv.close()
    h.free()


def test_server_shared_matrix(gpu):
    """Requests that share one matrix (the caller keeps it) give the same
    eigenvalues as a request that owns its copy, and the matrix stays usable."""
    tol = 1e-9
    (csr, eav), = _blocks([1500], 300, [0.05])
    want = _reference(csr, eav, 3)
    srv = rci.Server(8 << 30, max_active=8)
    h = rci.HamiltonianBlock.from_csr(csr, shift=-eav, precision=L.GCI_SPLIT32)
    rids = [srv.submit(h, k, tol, eav, shared=True) for k in (3, 2, 3)]
    srv.drain()
    res = [srv.poll(r)[0] for r in rids]
    srv.close()
    for got in res:
        k = len(got)
        assert np.all(np.abs(got - want[:k]) / np.abs(want[:k] + eav) < tol)
    # still the caller's: the operator runs on it after the server is gone
    B = np.asfortranarray(np.random.default_rng(4).standard_normal((csr.n, 2)))
    assert np.all(np.isfinite(h.spicmv2(B, L.GCI_OP_SPLIT_FULL)))
    h.free()


def test_batch_solve_result_buffers(gpu):
    """Eigenvectors copied into a caller's pageable buffer, a page-locked
    one (direct DMA) and a fresh array agree (runs differ only by the
    order of the SpMM's atomic sums)."""
    sizes, fills, nev = [700, 2600, 1500], [0.06, 0.03, 0.05], [3, 4, 2]
    data = _blocks(sizes, 700, fills)
    blocks = [rci.HamiltonianBlock.from_csr(c, shift=-e, precision=L.GCI_SPLIT32) for c, e in data]
    eav = [e for _, e in data]
    total = sum(n * k for n, k in zip(sizes, nev))
    ev0, V0, _ = rci.batch_solve(blocks, nev, 1e-9, eav)
    pinned = rci.pinned_empty(total + 5)
    ev1, V1, _ = rci.batch_solve(blocks, nev, 1e-9, eav, out=pinned)
    plain = np.empty(total)
    ev2, V2, _ = rci.batch_solve(blocks, nev, 1e-9, eav, out=plain)
    for b in range(len(blocks)):
        for ev, V in ((ev1, V1), (ev2, V2)):
            assert np.all(np.abs(ev[b] - ev0[b]) <= 1e-10 * np.abs(ev0[b] + eav[b]))
            assert np.all(np.abs(np.abs(np.sum(V[b] * V0[b], axis=0)) - 1.0) < 1e-8)
    assert np.shares_memory(V1[0], pinned) and np.shares_memory(V2[0], plain)
    with pytest.raises(ValueError):
        rci.batch_solve(blocks, nev, 1e-9, eav, out=np.empty(total - 1))


def test_batch_solve_decoupled_block(gpu):
    """A very sparse block whose Ritz vectors are exact eigenvectors: at the
    refinement restart A x lies in the basis span, which the Gram-Schmidt
    breakdown test must catch (a normalised remainder of rounding noise
    duplicated a basis vector and the solve never converged)."""
    from grasp_amd import synth
    h, eav = synth.synth_block(308, 1706206421809167660, fill=0.0107)
    iendc, irow, emt = h.download()
    # the held matrix (diagonal shifted by -EAV), densified by the oracle operator
    want = np.linalg.eigvalsh(O.spicmv2(iendc, irow, emt, np.eye(h.n)))
    for nev in (3, 5):
        ev, V, st = rci.batch_solve([h], [nev], 1e-8, [eav])
        assert st[0]["ierr"] == 0 and st[0]["nloops"] < 50, st[0]
        rel = np.abs(ev[0] - want[:nev]) / np.abs(want[:nev] + eav)
        assert np.all(rel < 1e-8), rel
    h.free()
