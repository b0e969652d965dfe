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


def _assert_residuals(csr, eav, V, ev, tol):
    """Eigenvector quality with the FP64 operator: ||H v - lambda v|| <=
    tol |lambda + EAV| per pair (the oracle's SPICMV2 on the held block,
    diagonal shifted by -EAV), and V^T V = I."""
    diag = np.zeros(csr.nnz, dtype=bool)
    diag[csr.row_end[1:] - 1] = True
    held = np.where(diag, csr.val - eav, csr.val)
    V = np.asfortranarray(V)
    r = np.linalg.norm(O.spicmv2(csr.row_end, csr.col, held, V, nthreads=8) - V * ev, axis=0)
    assert np.all(r <= tol * np.abs(ev + eav)), (r, tol * np.abs(ev + eav))
    k = V.shape[1]
    assert np.max(np.abs(V.T @ V - np.eye(k))) < 1e-8


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
        _assert_residuals(csr, eav, evecs[b], evals[b], tol)


@pytest.mark.parametrize("engines", [2, 3])
def test_batch_solve_engine_groups_identical(gpu, monkeypatch, engines):
    """GCI_BATCH_ENGINES splits the blocks into contiguous groups, one
    Davidson engine and stream each, driven in turn: every block's
    eigenpairs agree with the one-engine solve's to rounding (the SpMM's
    column sums arrive through atomics in no fixed order, and a group's
    launch width follows its own blocks)."""
    sizes = [120, 350, 900, 1600, 2400, 240, 3000, 700]
    fills = [0.1, 0.08, 0.05, 0.04, 0.03, 0.1, 0.02, 0.06]
    nev = [4, 2, 4, 3, 4, 1, 4, 2]
    data = _blocks(sizes, 300, fills)
    blocks = [rci.HamiltonianBlock.from_csr(c, shift=-e, precision=L.GCI_SPLIT32) for c, e in data]
    eavs = [e for _, e in data]
    monkeypatch.setenv("GCI_BATCH_ENGINES", "1")
    one = rci.batch_solve(blocks, nev, 1e-9, eavs)
    monkeypatch.setenv("GCI_BATCH_ENGINES", str(engines))
    many = rci.batch_solve(blocks, nev, 1e-9, eavs)
    for b in range(len(blocks)):
        rel = np.abs(one[0][b] - many[0][b]) / np.abs(one[0][b] + eavs[b])
        assert np.all(rel < 1e-12), (b, rel)
        overlap = np.abs(np.sum(one[1][b] * many[1][b], axis=0))
        assert np.all(np.abs(overlap - 1.0) < 1e-9), (b, overlap)
        assert many[2][b]["ierr"] == 0


def test_batch_solve_fp64_blocks(gpu):
    data = _blocks([400, 1500], 200, [0.1, 0.05])
    blocks = [rci.HamiltonianBlock.from_csr(c, shift=-e) for c, e in data]
    evals, evecs, stats = rci.batch_solve(blocks, [3, 3], 1e-10, [e for _, e in data])
    for b, (csr, eav) in enumerate(data):
        want = _reference(csr, eav, 3)
        assert np.all(np.abs(evals[b] - want) / np.abs(want + eav) < 1e-10)
        _assert_residuals(csr, eav, evecs[b], evals[b], 1e-10)


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
        alone.append(srv.poll(rid))
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
        for got in (alone[i], mix[i]):
            assert np.all(np.abs(got[0] - want) / np.abs(want + eav) < tol), i
            _assert_residuals(csr, eav, got[1], got[0], tol)
        assert np.all(np.abs(alone[i][0] - mix[i][0]) / np.abs(want + eav) < tol)
        # the same eigenvectors alone and in the mix (up to sign)
        ov = np.abs(np.sum(alone[i][1] * mix[i][1], axis=0))
        assert np.all(np.abs(ov - 1.0) < 1e-6), ov
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
    res = [srv.poll(r) for r in rids]
    srv.close()
    for got, vec, _ in res:
        k = len(got)
        assert np.all(np.abs(got - want[:k]) / np.abs(want[:k] + eav) < tol)
        _assert_residuals(csr, eav, vec, got, tol)
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


@pytest.mark.parametrize("engines", [1, 2])
def test_batch_solve_pinned_vectors(gpu, monkeypatch, engines):
    """A page-locked result buffer (direct copy) over blocks that finish at
    different iterations, with one and two engine groups: every block's
    vectors are its final ones (FP64-operator residuals within the
    tolerance, the pageable-buffer solve's vectors), a second solve into
    the same buffer overwrites it whole, and the guard elements past the
    end stay untouched."""
    sizes = [150, 3200, 400, 2600, 90, 1800]
    fills = [0.1, 0.02, 0.08, 0.03, 0.1, 0.04]
    nev = [2, 4, 3, 4, 1, 3]
    data = _blocks(sizes, 900, fills)
    blocks = [rci.HamiltonianBlock.from_csr(c, shift=-e, precision=L.GCI_SPLIT32) for c, e in data]
    eavs = [e for _, e in data]
    total = sum(n * k for n, k in zip(sizes, nev))
    monkeypatch.setenv("GCI_BATCH_ENGINES", str(engines))
    ev0, V0, _ = rci.batch_solve(blocks, nev, 1e-9, eavs, out=np.empty(total))
    pinned = rci.pinned_empty(total + 3)
    for rep in range(2):
        pinned[:] = np.nan
        pinned[total:] = 7.0
        ev1, V1, st = rci.batch_solve(blocks, nev, 1e-9, eavs, out=pinned)
        assert np.all(np.isfinite(pinned[:total])), rep
        assert np.all(pinned[total:] == 7.0)
        for b, (csr, eav) in enumerate(data):
            assert st[b]["ierr"] == 0
            _assert_residuals(csr, eav, V1[b], ev1[b], 1e-9)
            overlap = np.abs(np.sum(V0[b] * V1[b], axis=0))
            assert np.all(np.abs(overlap - 1.0) < 1e-8), (rep, b, overlap)


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
        Vb = np.asfortranarray(V[0])
        r = np.linalg.norm(O.spicmv2(iendc, irow, emt, Vb) - Vb * ev[0], axis=0)
        assert np.all(r <= 1e-8 * np.abs(ev[0] + eav)), r
    h.free()


def test_batch_unit_start_matches_spmm_start(gpu):
    """The batched solve's first product of its unit start vectors is read
    from H's columns (no SpMM pass); a solve of the same block whose first
    product goes through the SpMM (gci_davidson with the same unit vectors
    given as the initial estimates, same split operator and refinement)
    reaches the same eigenpairs."""
    (csr, eav), = _blocks([2600], 800, [0.04])
    nev, tol = 4, 1e-9
    h = rci.HamiltonianBlock.from_csr(csr, shift=-eav, precision=L.GCI_SPLIT32)
    ev_b, V_b, st_b = rci.batch_solve([h], [nev], tol, [eav])
    # unit vectors on the NEV smallest diagonal entries (ascending value, then index)
    d = csr.val[csr.row_end[1:] - 1]
    rows = np.lexsort((np.arange(csr.n), d))[:nev]
    init = np.zeros((csr.n, nev), order="F")
    init[rows, np.arange(nev)] = 1.0
    ev_g, V_g, res_g, st_g = rci.gdvd(h, nev, mblock=nev, init=init, mode=L.GCI_OP_SPLIT_BULK, refine=True)
    want = _reference(csr, eav, nev)
    for ev, V in ((ev_b[0], V_b[0]), (ev_g, V_g)):
        assert np.all(np.abs(ev - want) / np.abs(want + eav) < tol)
        _assert_residuals(csr, eav, V, ev, tol)
    assert np.all(np.abs(ev_b[0] - ev_g) / np.abs(want + eav) < tol)
    ov = np.abs(np.sum(V_b[0] * V_g, axis=0))
    assert np.all(np.abs(ov - 1.0) < 1e-6), ov
    h.free()


def test_server_refuses_bad_request_and_keeps_serving(gpu):
    """A request the engine cannot admit is refused at submit with
    GCI_ERR_ARG; requests around it (12 eigenpairs, LIM = 84, included)
    complete."""
    tol = 1e-9
    data = _blocks([2200, 40], 900, [0.04, 0.2])
    srv = rci.Server(8 << 30, max_active=4)
    (csr, eav), (small, seav) = data
    good = rci.HamiltonianBlock.from_csr(csr, shift=-eav, precision=L.GCI_SPLIT32)
    r12 = srv.submit(good, 12, tol, eav, shared=True)
    bad = rci.HamiltonianBlock.from_csr(small, shift=-seav, precision=L.GCI_SPLIT32)
    for nev in (40, 17):  # NEV = NCF (DVDSON IERR 128), NEV above 16
        with pytest.raises(L.GciError) as e:
            srv.submit(bad, nev, tol, seav, shared=True)
        assert e.value.code == L.GCI_ERR_ARG
    r2 = srv.submit(good, 2, tol, eav, shared=True)
    srv.drain()
    want = _reference(csr, eav, 12)
    for rid, k in ((r12, 12), (r2, 2)):
        ev, V, lat = srv.poll(rid)
        assert np.all(np.abs(ev - want[:k]) / np.abs(want[:k] + eav) < tol)
        _assert_residuals(csr, eav, V, ev, tol)
    srv.close()
    good.free()
    bad.free()
