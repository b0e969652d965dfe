"""Seeded synthetic inputs for the rci90 hot path.

Everything here is *input* generation: radial orbitals, CSF-block
Hamiltonians and angular-coefficient lists of the shapes the rci90 path
consumes.  Nothing here is measured and nothing here computes a result of
the path itself.

* Radial orbitals are screened-hydrogenic Dirac-Coulomb bound states for a
  point nucleus (closed form in confluent hypergeometric polynomials, the
  same physics rwfnestimate's screened-hydrogenic option produces --
  lib9290/dcbsrw.f90), tabulated on GRASP's exponential grid
  r_j = RNT (exp((j-1) H) - 1) (lib9290/radgrd.f90:41-60).
* Block Hamiltonians are symmetric, stored as lower-triangle rows with the
  diagonal last, exactly the .res row layout SETHAM writes
  (rci90/setham_gg.f90:378).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

C_LIGHT = 137.035999084  # speed of light in atomic units (CODATA 2018 1/alpha)

# ------------------------------------------------------------------ grid


def radgrd(n: int = 590, rnt: float = 2.0e-6, h: float = 0.05):
    """r, rp, rpor on n+10 points with the same recurrence as RADGRD."""
    r = np.zeros(n + 10)
    rp = np.zeros(n + 10)
    rpor = np.zeros(n + 10)
    rp[0] = rnt
    eph = math.exp(h)
    ett = 1.0
    for i in range(1, n + 10):
        ett = eph * ett
        ettm1 = ett - 1.0
        r[i] = rnt * ettm1
        rp[i] = rnt * ett
        rpor[i] = ett / ettm1
    return r, rp, rpor


# -------------------------------------------------------------- orbitals

_L_OF = {"s": 0, "p": 1, "d": 2, "f": 3, "g": 4, "h": 5, "i": 6}


def parse_subshell(name: str):
    """'2p-' -> (n=2, l=1, kappa=+1); '2p' -> (2, 1, -2)."""
    n = int("".join(ch for ch in name if ch.isdigit()))
    rest = name.lstrip("0123456789")
    l = _L_OF[rest[0]]
    minus = rest.endswith("-")
    kappa = l if minus else -(l + 1)
    return n, l, kappa


def subshells(nmax: int, lmax: int):
    """Relativistic subshell labels in GRASP order: n ascending, then l, j-."""
    out = []
    for n in range(1, nmax + 1):
        for l in range(0, min(n - 1, lmax) + 1):
            lab = "spdfghi"[l]
            if l > 0:
                out.append(f"{n}{lab}-")
            out.append(f"{n}{lab}")
    return out


def _poly_m(a: int, b: float, rho: np.ndarray) -> np.ndarray:
    """Kummer M(a, b, rho) for a = 0, -1, -2, ... (a terminating series)."""
    out = np.ones_like(rho)
    term = np.ones_like(rho)
    for s in range(0, -a):
        term = term * (a + s) / ((b + s) * (s + 1)) * rho
        out = out + term
    return out


def dirac_hydrogenic(n: int, kappa: int, z: float, r: np.ndarray, c: float = C_LIGHT):
    """Large/small radial functions P, Q (r times g, f) of the point-nucleus
    Dirac-Coulomb bound state (n, kappa) and its binding energy.  The
    normalisation is fixed numerically on the grid below (int P^2+Q^2 = 1)."""
    k = abs(kappa)
    nr = n - k
    za = z / c
    gam = math.sqrt(k * k - za * za)
    bign = math.sqrt(n * n - 2.0 * nr * (k - gam))
    eps = 1.0 / math.sqrt(1.0 + (za / (gam + nr)) ** 2)
    ebind = (1.0 - eps) * c * c
    rho = 2.0 * z * r / bign
    m1 = nr * _poly_m(-nr + 1, 2 * gam + 1, rho) if nr > 0 else np.zeros_like(rho)
    m0 = (bign - kappa) * _poly_m(-nr, 2 * gam + 1, rho)
    with np.errstate(under="ignore", over="ignore"):
        env = np.exp(-0.5 * rho) * np.power(rho, gam, where=rho > 0, out=np.zeros_like(rho))
    P = math.sqrt(1.0 + eps) * env * (m1 - m0)
    Q = math.sqrt(1.0 - eps) * env * (m1 + m0)
    return P, Q, ebind


@dataclass
class OrbitalSet:
    names: list
    P: np.ndarray     # (nw, n) large components on the grid
    Q: np.ndarray     # (nw, n) small components
    mf: np.ndarray    # last tabulated point per orbital (1-based, as MF)
    nak: np.ndarray   # kappa
    nkl: np.ndarray   # l
    nkj: np.ndarray   # 2j
    ebind: np.ndarray
    zeff: np.ndarray


def screened_orbitals(z: float, names, n: int = 590, h: float = 0.05, occupied=None,
                      c: float = C_LIGHT):
    """Screened-hydrogenic orbitals on the GRASP default grid for nuclear
    charge z: each subshell sees z_eff = z - sigma, sigma = electrons in
    lower-n shells of the reference configuration plus half of those in its
    own shell (occupied: dict n -> electrons)."""
    rnt = 2.0e-6 / z
    r, rp, _ = radgrd(n, rnt, h)
    r = r[:n]
    rp = rp[:n]
    occupied = occupied or {}
    accy = h ** 6
    nw = len(names)
    P = np.zeros((nw, n))
    Q = np.zeros((nw, n))
    mf = np.zeros(nw, dtype=np.int32)
    nak = np.zeros(nw, dtype=np.int32)
    nkl = np.zeros(nw, dtype=np.int32)
    nkj = np.zeros(nw, dtype=np.int32)
    eb = np.zeros(nw)
    zeff = np.zeros(nw)
    for j, name in enumerate(names):
        nn, l, kappa = parse_subshell(name)
        sigma = sum(v for m, v in occupied.items() if m < nn) + 0.5 * max(occupied.get(nn, 0) - 1, 0)
        ze = max(z - sigma, 1.0)
        p, q, e = dirac_hydrogenic(nn, kappa, ze, r, c)
        # normalise on the grid: int (P^2 + Q^2) dr with dr = rp dt
        w = np.full(n, h)
        w[0] = 0.5 * h
        w[-1] = 0.5 * h
        norm = math.sqrt(float(np.sum((p * p + q * q) * rp * w)))
        p /= norm
        q /= norm
        # sign: P positive near the origin
        i0 = int(np.argmax(np.abs(p) > 0))
        if p[i0] < 0:
            p, q = -p, -q
        # MF: last point above ACCY/10 (the tabulation cut of lib9290/dcbsrw.f90)
        cut = 0.1 * accy
        big = np.nonzero(np.abs(p) >= cut)[0]
        m = int(big[-1]) + 1 if len(big) else n
        m = min(m, n - 4)
        p[m:] = 0.0
        q[m:] = 0.0
        P[j], Q[j] = p, q
        mf[j] = m
        nak[j] = kappa
        nkl[j] = l
        nkj[j] = 2 * abs(kappa) - 1
        eb[j] = e
        zeff[j] = ze
    return OrbitalSet(list(names), P, Q, mf, nak, nkl, nkj, eb, zeff)


def fermi_parameters(a_mass: float):
    """Fermi nucleus (c, a) in bohr from the usual R_rms = 0.836 A^(1/3) + 0.570
    fm and skin thickness t = 2.3 fm, a = t / (4 ln 3)."""
    fm = 1.0e-15 / 0.529177210903e-10
    rrms = 0.836 * a_mass ** (1.0 / 3.0) + 0.570
    a = 2.3 / (4.0 * math.log(3.0))
    c2 = (5.0 / 3.0) * rrms ** 2 - (7.0 / 3.0) * (math.pi * a) ** 2
    return math.sqrt(c2) * fm, a * fm


# --------------------------------------------------- block Hamiltonians


@dataclass
class LowerCSR:
    """Lower-triangle rows, diagonal last, 1-based columns (the .res layout)."""
    n: int
    row_end: np.ndarray   # int64, n+1, row_end[0] = 0
    col: np.ndarray       # int32, 1-based
    val: np.ndarray       # float64

    @property
    def nnz(self) -> int:
        return int(self.row_end[-1])

    def dense(self) -> np.ndarray:
        a = np.zeros((self.n, self.n))
        rows = np.repeat(np.arange(self.n), np.diff(self.row_end))
        cols = self.col - 1
        a[rows, cols] = self.val
        a[cols, rows] = self.val
        return a


def clustered_block(n: int, seed: int, fill: float = 0.05, diag_mean: float = -338.0,
                    diag_sd: float = 5.0, off_sd: float = 1.0e-2, group: tuple = (4, 48),
                    inner: float = 0.6, sorted_diag: bool = False) -> LowerCSR:
    """Seeded symmetric block: diagonal ~ N(diag_mean, diag_sd^2), off-diagonal
    entries ~ N(0, off_sd^2) clustered in CSF-group sub-blocks.  Rows are cut
    into groups of uniform random size in [group[0], group[1]]; a lower group
    pair (G, H) is coupled with probability fill/inner and, when coupled,
    each entry is present with probability inner, so the expected lower fill
    is `fill`.  sorted_diag puts the diagonal in ascending order, as a CSF
    list does when it starts from the reference configuration (so the
    leading-block initial estimates of INIEST2 see the lowest states)."""
    rng = np.random.default_rng(seed)
    bounds = [0]
    while bounds[-1] < n:
        bounds.append(min(n, bounds[-1] + int(rng.integers(group[0], group[1] + 1))))
    bounds = np.array(bounds)
    ng = len(bounds) - 1
    pcoup = min(1.0, fill / inner)
    rows_c = []
    cols_c = []
    for G in range(ng):
        r0, r1 = bounds[G], bounds[G + 1]
        Hs = np.nonzero(rng.random(G + 1) < pcoup)[0]
        for H in Hs:
            c0, c1 = bounds[H], bounds[H + 1]
            m = rng.random((r1 - r0, c1 - c0)) < inner
            if H == G:
                m = np.tril(m, -1)
            ri, ci = np.nonzero(m)
            rows_c.append(ri + r0)
            cols_c.append(ci + c0)
    if rows_c:
        rows = np.concatenate(rows_c)
        cols = np.concatenate(cols_c)
    else:
        rows = np.zeros(0, dtype=np.int64)
        cols = np.zeros(0, dtype=np.int64)
    vals = rng.normal(0.0, off_sd, size=len(rows))
    # sort by (row, col), then append the diagonal last in each row
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    diag = rng.normal(diag_mean, diag_sd, size=n)
    if sorted_diag:
        diag = np.sort(diag)
    counts = np.bincount(rows, minlength=n) + 1
    row_end = np.zeros(n + 1, dtype=np.int64)
    row_end[1:] = np.cumsum(counts)
    nnz = int(row_end[-1])
    col = np.zeros(nnz, dtype=np.int32)
    val = np.zeros(nnz)
    # position of each off-diagonal entry: row start + rank within row
    start = row_end[:-1]
    rank = np.arange(len(rows)) - np.searchsorted(rows, rows, side="left")
    pos_off = start[rows] + rank
    col[pos_off] = cols + 1
    val[pos_off] = vals
    dpos = row_end[1:] - 1
    col[dpos] = np.arange(1, n + 1)
    val[dpos] = diag
    return LowerCSR(n, row_end, col, val)


def batch_sizes(nblocks: int, seed: int, lo: float = 1.0e2, hi: float = 5.0e4):
    """Block sizes log-uniform in [lo, hi] (seeded)."""
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(math.log(lo), math.log(hi), size=nblocks))).astype(np.int64)


# ------------------------------------------------------ angular lists


def synthetic_lists(csr_pattern: LowerCSR, nw: int, nkl, rk_labels, rk_k, seed: int,
                    max_terms: int = 6, small_frac: float = 0.1):
    """Angular-coefficient lists in the layout the CPU producer emits per CSF
    pair (ONESCALAR's IA, IB, TSHELL; RKCO_GG's LABEL(1:5,i), COEFF(i)),
    on the sparsity pattern of csr_pattern (pairs = its stored elements).

    Two-body labels are drawn from the R^k table and permuted by the
    symmetries RKINTC canonicalises ((a,b,c,d) -> (c,b,a,d), (a,d,c,b),
    (c,d,a,b)); a fraction of coefficients is put below the 1e-12 CUTOFF.
    Diagonal pairs carry IA = IB = 1 with NW TSHELL values (occupations);
    off-diagonal pairs carry a same-kappa IA != IB pair or IA = 0."""
    rng = np.random.default_rng(seed)
    n = csr_pattern.n
    rows = np.repeat(np.arange(1, n + 1), np.diff(csr_pattern.row_end)).astype(np.int32)
    cols = csr_pattern.col.astype(np.int32)
    npairs = len(rows)
    nkl = np.asarray(nkl)
    # same-l pairs a < b (kappa equal is enough; l equal is a superset check)
    same = [(a, b) for a in range(1, nw + 1) for b in range(a + 1, nw + 1) if nkl[a - 1] == nkl[b - 1]]
    one_ia = np.zeros(npairs, dtype=np.int32)
    one_ib = np.zeros(npairs, dtype=np.int32)
    diag = rows == cols
    choose = rng.random(npairs)
    sel = (~diag) & (choose < 0.3) if same else np.zeros(npairs, dtype=bool)
    # TSHELL values per pair: NW on the diagonal, one for a selected
    # off-diagonal one-body pair, none otherwise
    ntsh = np.where(diag, nw, np.where(sel, 1, 0)).astype(np.int64)
    tsh_off = (np.cumsum(ntsh) - ntsh).astype(np.int64)
    tshell = np.zeros(max(1, int(ntsh.sum())))
    di = np.nonzero(diag)[0]
    one_ia[di] = one_ib[di] = 1
    occ = rng.integers(0, 3, size=(len(di), nw)).astype(np.float64)
    occ[rng.random(occ.shape) < 0.1] = 1e-13
    tshell[(tsh_off[di][:, None] + np.arange(nw)[None, :]).ravel()] = occ.ravel()
    si = np.nonzero(sel)[0]
    if len(si):
        pairs = np.asarray(same, dtype=np.int32)[rng.integers(len(same), size=len(si))]
        swap = rng.random(len(si)) < 0.5
        one_ia[si] = np.where(swap, pairs[:, 1], pairs[:, 0])
        one_ib[si] = np.where(swap, pairs[:, 0], pairs[:, 1])
        tshell[tsh_off[si]] = rng.normal(0, 0.3, size=len(si))
    nt = rng.integers(1, max_terms + 1, size=npairs)
    tb_end = np.cumsum(nt).astype(np.int64)
    tb_beg = (tb_end - nt).astype(np.int64)
    nterms = int(tb_end[-1])
    pick = rng.integers(0, len(rk_k), size=nterms)
    lab = np.asarray(rk_labels)[pick].copy()
    kk = np.asarray(rk_k)[pick]
    sym = rng.integers(0, 4, size=nterms)
    a, b, c, d = lab[:, 0].copy(), lab[:, 1].copy(), lab[:, 2].copy(), lab[:, 3].copy()
    m1 = sym == 1
    lab[m1, 0], lab[m1, 2] = c[m1], a[m1]
    m2 = sym == 2
    lab[m2, 1], lab[m2, 3] = d[m2], b[m2]
    m3 = sym == 3
    lab[m3, 0], lab[m3, 1], lab[m3, 2], lab[m3, 3] = c[m3], d[m3], a[m3], b[m3]
    label = np.concatenate([lab, kk[:, None]], axis=1).astype(np.int32).ravel()
    coeff = rng.normal(0.0, 0.1, size=nterms)
    coeff[rng.random(nterms) < small_frac] = 1e-13
    return {
        "pair_ic": rows, "pair_ir": cols, "one_ia": one_ia, "one_ib": one_ib, "tsh_off": tsh_off,
        "tshell": tshell, "tb_beg": tb_beg, "tb_end": tb_end, "label": label, "coeff": coeff,
    }
