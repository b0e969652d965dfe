"""Pins of the CPU oracle (oracle/graspci_oracle.c) -- no GPU needed.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong index or sign) fails it:
* the reference's own known answer (test/lib9290_quad.f90:22-34);
* Newton-Cotes / Lagrange tables against polynomial exactness;
* closed forms for hydrogenic Slater integrals and Dirac energies;
* symmetries of R^k, and dense linear algebra from numpy/scipy;
* the compiled reference itself (oracle/_ref, when it was built here).
"""
import math

import numpy as np
import pytest

import oracle as O
from grasp_amd import workloads as W


@pytest.fixture(scope="module")
def grid():
    return O.Grid(590, 2.0e-6, 0.05)


def test_quad_reference_known_answer(grid):
    # test/lib9290_quad.f90: QUAD of exp(-2r) on the default grid = 1/2 within 1e-12
    ta = np.exp(-2.0 * grid.r) * grid.rp
    assert abs(grid.quad(ta, 590) - 0.5) <= 1e-12


def test_newton_cotes_tables_exact_for_polynomials(grid):
    # closed 5-point rule integrates degree <= 5 exactly; the starting
    # formulas (integral from node 1 to node j) degree <= 4
    h = grid.h
    c1, c2, c3, c4 = grid.ncc
    f = lambda t: 1 + 2 * t - 3 * t ** 2 + 0.5 * t ** 3 - 0.25 * t ** 4 + 0.125 * t ** 5
    F = lambda t: t + t ** 2 - t ** 3 + 0.125 * t ** 4 - 0.05 * t ** 5 + 0.125 / 6 * t ** 6
    x = np.arange(5) * h
    assert abs((c1 * (f(x[0]) + f(x[4])) + c2 * (f(x[1]) + f(x[3])) + c3 * f(x[2])) - (F(4 * h) - F(0))) < 1e-14
    assert abs(c4 - 2 * c1) == 0.0
    cnc = grid.cnc5c.reshape(3, 4)  # [j-2][i-2]
    g = lambda t: 3 - t + 2 * t ** 2 - t ** 3 + 0.5 * t ** 4
    G = lambda t: 3 * t - t ** 2 / 2 + 2 * t ** 3 / 3 - t ** 4 / 4 + 0.1 * t ** 5
    for j in range(2, 5):
        # node 1 (t=0) carries the omitted weight 251/232/243 * h/720 etc.
        w1 = {2: 251, 3: 232, 4: 243}[j] * h / 720
        approx = w1 * g(0) + sum(cnc[j - 2][i - 2] * g((i - 1) * h) for i in range(2, 6))
        assert abs(approx - (G((j - 1) * h) - G(0))) < 1e-15


def test_lagrange13_derivative_exact(grid):
    a = grid.a13.reshape(13, 13)
    t = np.arange(1, 14, dtype=float)
    for deg in range(0, 13):
        f = t ** deg
        df = deg * t ** (deg - 1) if deg else np.zeros(13)
        assert np.max(np.abs(a @ f - df)) <= 1e-9 * max(1.0, np.max(np.abs(df)))


def test_radgrd_closed_form():
    g = O.Grid(590, 2e-6 / 26, 0.05)
    j = np.arange(600)
    assert np.allclose(g.r, 2e-6 / 26 * (np.exp(j * 0.05) - 1), rtol=1e-13, atol=0)
    assert np.allclose(g.rp, 2e-6 / 26 * np.exp(j * 0.05), rtol=1e-13)


def _hydrogenic_set(z, names, n=590):
    orb = W.screened_orbitals(z, names, n=n, occupied={})  # unscreened: z_eff = z
    return orb, O.Orbitals(orb.P, orb.Q, orb.mf, orb.nak, orb.nkl, orb.nkj)


def test_f0_1s1s_nonrelativistic_limit():
    # F0(1s,1s) = 5Z/8 for hydrogenic 1s; Dirac corrections O((Z alpha)^2)
    z = 2.0
    g = O.Grid(590, 2e-6 / z, 0.05)
    orb, o = _hydrogenic_set(z, ["1s"])
    f0 = O.slater(g, o, 1, 1, 1, 1, 0)
    assert abs(f0 - 5 * z / 8) / (5 * z / 8) < 5e-4


def test_rinti_dirac_energy_point_nucleus():
    # I(a,a) of an exact point-nucleus Dirac-Coulomb orbital = -(binding energy)
    z = 26.0
    g = O.Grid(590, 2e-6 / z, 0.05)
    names = ["1s", "2s", "2p-", "2p", "3d"]
    orb, o = _hydrogenic_set(z, names)
    zz = O.nucpot(g, z)
    for j in range(1, len(names) + 1):
        val = O.rinti(g, o, zz, W.C_LIGHT, j, j, 0)
        assert abs(val + orb.ebind[j - 1]) / orb.ebind[j - 1] < 1e-7, names[j - 1]


def test_rk_symmetries():
    z = 6.0
    g = O.Grid(590, 2e-6 / z, 0.05)
    orb = W.screened_orbitals(z, ["1s", "2s", "2p-", "2p", "3s", "3d"], occupied={1: 2, 2: 2})
    o = O.Orbitals(orb.P, orb.Q, orb.mf, orb.nak, orb.nkl, orb.nkj)
    # R^k(ab;cd) = R^k(cb;ad) = R^k(ad;cb) exactly: the swaps stay inside one
    # density product, so SLATER's arithmetic is unchanged (slater.f90:288-291,
    # yzk.f90:107-109).  R^k(ba;dc) exchanges which density YZK integrates, so
    # it agrees only to quadrature accuracy.
    for (a, b, c, d, k) in [(1, 2, 1, 2, 0), (1, 4, 3, 5, 1), (2, 6, 5, 2, 2)]:
        v = O.slater(g, o, a, b, c, d, k)
        for (p, q, r, s) in [(c, b, a, d), (a, d, c, b)]:
            assert O.slater(g, o, p, q, r, s, k) == v
        assert abs(O.slater(g, o, b, a, d, c, k) - v) < 1e-4


def test_genintrk_keys_sorted_and_lookup():
    z = 6.0
    g = O.Grid(590, 2e-6 / z, 0.05)
    names = W.subshells(3, 2)
    orb = W.screened_orbitals(z, names, occupied={1: 2, 2: 2})
    o = O.Orbitals(orb.P, orb.Q, orb.mf, orb.nak, orb.nkl, orb.nkj)
    kstart, keys, vals, kv, labels = O.genintrk(g, o)
    for k in range(len(kstart) - 1):
        seg = keys[kstart[k] - 1: kstart[k + 1] - 1]
        assert np.all(np.diff(seg) > 0)
    # RKINTC canonicalisation: every symmetric variant finds the same value
    rng = np.random.default_rng(0)
    for t in rng.choice(len(keys), 40, replace=False):
        a, b, c, d = labels[t]
        k = kv[t]
        for (p, q, r, s) in [(a, b, c, d), (c, b, a, d), (a, d, c, b), (c, d, a, b)]:
            assert O.rkintc(o.nw, kstart, keys, vals, p, q, r, s, k) == vals[t]


def test_spicmv2_matches_dense():
    csr = W.clustered_block(300, 1, fill=0.1)
    B = np.random.default_rng(1).standard_normal((300, 3))
    got = O.spicmv2(csr.row_end, csr.col, csr.val, B)
    assert np.allclose(got, csr.dense() @ B, rtol=1e-13, atol=1e-12)
    got_mt = O.spicmv2(csr.row_end, csr.col, csr.val, B, nthreads=4)
    assert np.allclose(got_mt, got, rtol=1e-13, atol=1e-12)


def test_syev_matches_numpy():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((60, 60))
    a = a + a.T
    w, v = O.syev(a)
    assert np.allclose(w, np.linalg.eigvalsh(a), rtol=1e-12, atol=1e-12)
    assert np.allclose(v.T @ a @ v, np.diag(w), atol=1e-10)


@pytest.mark.parametrize("kind", ["random", "clustered", "degenerate", "diagonal", "block"])
def test_spevx_lowest_matches_numpy(kind):
    """The DSPEVX-style selected solver against numpy's eigh (LAPACK
    dsyevd), on spectra that stress each step: random, a tight cluster at
    the bottom (inverse iteration's reorthogonalisation), exactly repeated
    eigenvalues, a diagonal matrix (every reflector is the identity) and a
    decoupled block diagonal matrix (zero off-diagonals in T)."""
    rng = np.random.default_rng(7)
    n, nev = 257, 12
    if kind == "random":
        a = rng.standard_normal((n, n))
        a = a + a.T
    else:
        q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        lam = np.sort(rng.uniform(-5, 5, n))
        if kind == "clustered":
            lam[:6] = -10.0 + 1e-9 * np.arange(6)
        elif kind == "degenerate":
            lam[:4] = -10.0
            lam[4:7] = -9.0
        if kind == "diagonal":
            a = np.diag(rng.permutation(lam))
        elif kind == "block":
            b = rng.standard_normal((n // 2, n // 2))
            a = np.zeros((n, n))
            a[: n // 2, : n // 2] = b + b.T
            a[n // 2:, n // 2:] = np.diag(lam[n // 2:])
        else:
            a = (q * lam) @ q.T
            a = 0.5 * (a + a.T)
    want = np.linalg.eigvalsh(a)[:nev]
    for threads in (1, 4):
        w, z = O.spevx_lowest(a, nev, nthreads=threads)
        scale = np.max(np.abs(np.linalg.eigvalsh(a)))
        assert np.max(np.abs(w - want)) <= 1e-13 * scale, kind
        # eigenvectors: residuals and orthonormality (bases of repeated
        # eigenvalues are not unique)
        assert np.max(np.abs(a @ z - z * w)) <= 1e-11 * scale, kind
        assert np.max(np.abs(z.T @ z - np.eye(nev))) <= 1e-11, kind


def test_spevx_lowest_matches_syev():
    csr = W.clustered_block(400, 6, fill=0.1)
    a = csr.dense()
    w1, z1 = O.spevx_lowest(a, 5)
    w2, z2 = O.syev(a)
    assert np.allclose(w1, w2[:5], rtol=0, atol=1e-12 * np.max(np.abs(w2)))
    assert np.allclose(np.abs(np.sum(z1 * z2[:, :5], axis=0)), 1.0, atol=1e-9)


def test_gdvd_matches_eigh():
    n, nvec = 400, 3
    csr = W.clustered_block(n, 3, fill=0.1, sorted_diag=True)
    dense = csr.dense()
    d = dense.diagonal().copy()
    init = O.iniest2(csr.row_end, csr.col, csr.val, nvec, nmax=120)
    ev, vec, nl, nm, ierr, res = O.gdvd(csr.row_end, csr.col, csr.val, d, nvec, 2 * nvec + 60, init)
    assert ierr == 0
    w = np.linalg.eigvalsh(dense)[:nvec]
    assert np.allclose(ev, w, rtol=0, atol=1e-9)
    # TSTSEL may stop on CRITC/CRITE before CRITR (dvdson.f90:1328-1370)
    assert np.all(res < 1e-6)
    assert np.max(np.abs(vec.T @ vec - np.eye(nvec))) < 1e-10


def test_gdvd_block_mblock():
    n, nvec = 300, 4
    csr = W.clustered_block(n, 4, fill=0.15, sorted_diag=True)
    dense = csr.dense()
    init = O.iniest2(csr.row_end, csr.col, csr.val, nvec, nmax=50)
    ev, vec, nl, nm, ierr, res = O.gdvd(csr.row_end, csr.col, csr.val, dense.diagonal().copy(), nvec, 68, init,
                                        mblock=nvec)
    assert ierr == 0
    assert np.allclose(ev, np.linalg.eigvalsh(dense)[:nvec], atol=1e-9)


def test_maneig_dense_and_davidson_agree():
    csr = W.clustered_block(2100, 5, fill=0.04, sorted_diag=True)
    eav = float(np.mean(csr.val[csr.row_end[1:] - 1]))
    ev, vec, nl, nm, rc = O.maneig(csr.row_end, csr.col, csr.val, eav, 2, nthreads=4)
    assert rc == 0 and nl > 0
    w = np.linalg.eigvalsh(csr.dense() - eav * np.eye(csr.n))[:2]
    assert np.allclose(ev, w, atol=1e-9)
    # sign convention: dominant component positive
    for j in range(2):
        assert vec[np.argmax(np.abs(vec[:, j])), j] > 0


def test_maneig_mcdhf_levels_and_branches():
    """rmcdhf90's MANEIG: dense INIEST2 up to 4000 CSFs (levels ICCMIN picked
    out of the lowest NVECT, skipping one), and the warm-started GDVD above
    4000, against the library eigensolvers (numpy eigh, scipy eigsh)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.linalg import eigsh

    csr = W.clustered_block(600, 21, fill=0.06, sorted_diag=True)
    eav = float(np.mean(csr.val[csr.row_end[1:] - 1]))
    ev, vec, nl, nm, rc = O.maneig_mcdhf(csr.row_end, csr.col, csr.val, eav, 4, [2, 4], nthreads=4)
    w, z = np.linalg.eigh(csr.dense() - eav * np.eye(csr.n))
    assert rc == 0 and nl == 0
    assert np.allclose(ev, w[[1, 3]], atol=1e-10)
    for j, lev in enumerate([1, 3]):
        assert abs(abs(vec[:, j] @ z[:, lev]) - 1.0) < 1e-9
        assert vec[np.argmax(np.abs(vec[:, j])), j] > 0
    # warm start above 4000 CSFs from the previous SCF iteration's vectors
    # (the same block with off-diagonals scaled by 0.99); level 2 starts
    # from its unit vector
    n = 4200
    big = W.clustered_block(n, 22, fill=0.02, sorted_diag=True)
    eav = float(np.mean(big.val[big.row_end[1:] - 1]))
    rows = np.repeat(np.arange(n), np.diff(big.row_end))
    cols = big.col.astype(np.int64) - 1
    diag_pos = big.row_end[1:] - 1

    def lowest(scale):
        vals = big.val * scale
        vals[diag_pos] = big.val[diag_pos] - eav
        A = csr_matrix((vals, (rows, cols)), shape=(n, n))
        A = A + A.T - csr_matrix((vals[diag_pos], (np.arange(n), np.arange(n))), shape=(n, n))
        w, z = eigsh(A, k=3, which="SA", tol=0)
        order = np.argsort(w)
        return w[order], z[:, order]

    w0, z0 = lowest(0.99)
    w, z = lowest(1.0)
    ev, vec, nl, nm, rc = O.maneig_mcdhf(big.row_end, big.col, big.val, eav, 3, [1, 3],
                                         warm=(w0[[0, 2]], z0[:, [0, 2]]), nthreads=4)
    assert rc == 0 and nl > 0
    assert np.allclose(ev, w[[0, 2]], atol=1e-9)


def test_setham_resums_lists_in_order():
    # two pairs, hand-made lists: diagonal (1,1) and (2,1), (2,2)
    nw = 2
    iab = np.array([[1.5, 0.25], [0.0, -0.75]])
    kstart = np.array([1, 2, 2], dtype=np.int32)
    keys = np.array([((1 * 3 + 1) * 3 + 1) * 3 + 1], dtype=np.int32)
    vals = np.array([0.625])
    lists = {
        "pair_ic": np.array([1, 2, 2], dtype=np.int32), "pair_ir": np.array([1, 1, 2], dtype=np.int32),
        "one_ia": np.array([1, 1, 2], dtype=np.int32), "one_ib": np.array([1, 2, 2], dtype=np.int32),
        "tsh_off": np.array([0, 2, 3], dtype=np.int64),
        "tshell": np.array([2.0, 1e-13, 4.0, 1.0, 1.0]),
        "tb_beg": np.array([0, 1, 1], dtype=np.int64), "tb_end": np.array([1, 1, 2], dtype=np.int64),
        "label": np.array([1, 1, 1, 1, 0, 1, 1, 1, 1, 0], dtype=np.int32),
        "coeff": np.array([0.5, 2.0]),
    }
    row_end, col, val, eav = O.setham(2, nw, lists, iab, kstart, keys, vals)
    # (1,1): 2*I11 (+1e-13 dropped by CUTOFF) + 0.5*R0(11;11)
    assert val[0] == 2.0 * 1.5 + 0.5 * 0.625
    # (2,1): IA != IB -> TSHELL(1)*I(1,2)
    assert val[1] == 4.0 * 0.25
    # (2,2): IA == IB -> sum over orbitals of TSHELL(a) I(a,a), then 2*R0
    assert val[2] == (1.0 * 1.5 + 1.0 * -0.75) + 2.0 * 0.625
    assert list(row_end) == [0, 1, 3] and list(col) == [1, 1, 2]
    assert eav == val[0] + val[2]
