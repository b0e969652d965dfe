"""CPU oracle for the rci90 hot path -- TEST INFRASTRUCTURE ONLY.

Thin ctypes bindings over ``liboracle.so`` (oracle/graspci_oracle.c), the C
restatement of the GRASP2018 routines on the path.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package; the product package ``grasp_amd`` never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_int = C.POINTER(C.c_int)


def build() -> str:
    """Compile liboracle.so in place (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_quad.restype = C.c_double
        L.orc_slater.restype = C.c_double
        L.orc_rinti.restype = C.c_double
        L.orc_yzk.restype = C.c_int
        L.orc_genintrk_enum.restype = C.c_int64
        L.orc_setham.restype = C.c_int64
        L.orc_gdvd_worksize.restype = C.c_int64
        L.orc_gdvd_worksize.argtypes = [C.c_int, C.c_int, C.c_int]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def dp(a):
    return _p(a, _d)


def i32p(a):
    return _p(a, _i32)


def i64p(a):
    return _p(a, _i64)


def ip(a):
    return _p(a, _int)


# ----------------------------------------------------------------- grid

class Grid:
    """RADGRD + SETQIC state (lib9290/radgrd.f90, lib9290/setqic.f90)."""

    def __init__(self, n: int = 590, rnt: float = 2.0e-6, h: float = 0.05):
        self.n, self.rnt, self.h = n, rnt, h
        L = lib()
        self.r = np.zeros(n + 10)
        self.rp = np.zeros(n + 10)
        self.rpor = np.zeros(n + 10)
        L.orc_radgrd(C.c_int(n), C.c_double(rnt), C.c_double(h), dp(self.r), dp(self.rp), dp(self.rpor))
        self.ncc = np.zeros(4)
        self.cnc5c = np.zeros(12)
        L.orc_setqic(C.c_double(h), dp(self.ncc), dp(self.cnc5c))
        self.a13 = np.zeros(169)
        L.orc_lagrange13(dp(self.a13))
        self.accy = h ** 6  # rci90/getcid.f90:262

    def quad(self, ta: np.ndarray, mtp: int) -> float:
        buf = np.zeros(max(len(ta), mtp + 3) + 8)
        buf[: len(ta)] = ta
        return lib().orc_quad(dp(self.r), dp(self.rp), dp(self.ncc), dp(buf), C.c_int(mtp))


class Orbitals:
    """Radial functions P, Q (nw x ld), MF, kappa, l, 2j per orbital."""

    def __init__(self, P, Q, mf, nak, nkl, nkj):
        self.P = np.ascontiguousarray(P, dtype=np.float64)
        self.Q = np.ascontiguousarray(Q, dtype=np.float64)
        self.nw, self.ld = self.P.shape
        self.mf = np.ascontiguousarray(mf, dtype=np.int32)
        self.nak = np.ascontiguousarray(nak, dtype=np.int32)
        self.nkl = np.ascontiguousarray(nkl, dtype=np.int32)
        self.nkj = np.ascontiguousarray(nkj, dtype=np.int32)


def yzk(g: Grid, o: Orbitals, k: int, i: int, j: int):
    """Y^k(i,j;r) for 0-based orbitals; returns (yk[n], mtp)."""
    yk = np.zeros(g.n + 16)
    mtp = lib().orc_yzk(C.c_int(g.n), dp(g.r), dp(g.rp), dp(g.ncc), dp(g.cnc5c), C.c_double(g.accy),
                        C.c_int(k), dp(o.P), dp(o.Q), C.c_int(o.ld), i32p(o.mf), C.c_int(i), C.c_int(j), dp(yk))
    return yk[: g.n], mtp


def slater(g: Grid, o: Orbitals, ia, ib, ic, id_, k) -> float:
    """R^k(ia ib; ic id), 1-based labels."""
    return lib().orc_slater(C.c_int(g.n), dp(g.r), dp(g.rp), dp(g.rpor), dp(g.ncc), dp(g.cnc5c),
                            C.c_double(g.accy), dp(o.P), dp(o.Q), C.c_int(o.ld), i32p(o.mf),
                            C.c_int(ia), C.c_int(ib), C.c_int(ic), C.c_int(id_), C.c_int(k))


def nucpot(g: Grid, z: float, parm=None) -> np.ndarray:
    zz = np.zeros(g.n)
    if parm is None:
        lib().orc_nucpot(C.c_int(g.n), dp(g.r), C.c_double(z), C.c_int(0), None, dp(zz))
    else:
        p = np.asarray(parm, dtype=np.float64)
        lib().orc_nucpot(C.c_int(g.n), dp(g.r), C.c_double(z), C.c_int(2), dp(p), dp(zz))
    return zz


def rinti(g: Grid, o: Orbitals, zz: np.ndarray, c: float, j: int, k: int, mode: int = 0) -> float:
    return lib().orc_rinti(C.c_int(g.n), dp(g.r), dp(g.rp), dp(g.rpor), dp(g.ncc), dp(g.a13),
                           C.c_double(g.h), C.c_double(c), dp(zz), dp(o.P), dp(o.Q), C.c_int(o.ld),
                           i32p(o.mf), i32p(o.nak), C.c_int(j), C.c_int(k), C.c_int(mode))


def genintrk_enum(o: Orbitals):
    """GENINTRK enumeration: (kstart[0..j2max+1], keys, k, labels[n,4])."""
    L = lib()
    j2 = C.c_int(0)
    n = L.orc_genintrk_enum(C.c_int(o.nw), i32p(o.nkl), i32p(o.nkj), C.byref(j2), None, None, None, None)
    kstart = np.zeros(j2.value + 2, dtype=np.int32)
    keys = np.zeros(max(n, 1), dtype=np.int32)
    kv = np.zeros(max(n, 1), dtype=np.int32)
    labels = np.zeros((max(n, 1), 4), dtype=np.int32)
    L.orc_genintrk_enum(C.c_int(o.nw), i32p(o.nkl), i32p(o.nkj), C.byref(j2), i32p(kstart), i32p(keys),
                        i32p(kv), i32p(labels))
    return kstart, keys[:n], kv[:n], labels[:n]


def genintrk(g: Grid, o: Orbitals):
    """The full R^k table (keys, values, KSTART) as GENINTRK builds it."""
    kstart, keys, kv, labels = genintrk_enum(o)
    vals = np.array([slater(g, o, int(a), int(b), int(c), int(d), int(k))
                     for (a, b, c, d), k in zip(labels, kv)], dtype=np.float64)
    return kstart, keys, vals, kv, labels


def rkintc(nw, kstart, keys, vals, ia, ib, ic, id_, k):
    v = C.c_double(0.0)
    rc = lib().orc_rkintc(C.c_int(nw), i32p(kstart), i32p(keys), dp(vals), C.c_int(ia), C.c_int(ib),
                          C.c_int(ic), C.c_int(id_), C.c_int(k), C.byref(v))
    if rc:
        raise KeyError("Rkintc Integral not found")
    return v.value


def setham(ncf, nw, lists, iab, kstart, keys, vals, lfordr=False, iccut=0):
    """SETHAM reduction over angular lists; returns (row_end, col, val, eav_sum)."""
    L = lib()
    npairs = len(lists["pair_ic"])
    cap = npairs + 1
    row_end = np.zeros(ncf + 1, dtype=np.int64)
    col = np.zeros(cap, dtype=np.int32)
    val = np.zeros(cap)
    eav = C.c_double(0.0)
    nnz = L.orc_setham(C.c_int(ncf), C.c_int(nw), C.c_int64(npairs),
                       i32p(lists["pair_ic"]), i32p(lists["pair_ir"]), i32p(lists["one_ia"]),
                       i32p(lists["one_ib"]), i64p(lists["tsh_off"]), dp(lists["tshell"]),
                       i64p(lists["tb_beg"]), i64p(lists["tb_end"]), i32p(lists["label"]),
                       dp(lists["coeff"]), dp(np.ascontiguousarray(iab, dtype=np.float64)),
                       i32p(kstart), i32p(keys), dp(vals), C.c_int(int(lfordr)), C.c_int(iccut),
                       i64p(row_end), i32p(col), dp(val), C.byref(eav))
    if nnz < 0:
        raise ValueError("setham: malformed list or missing integral")
    return row_end, col[:nnz].copy(), val[:nnz].copy(), eav.value


def setham_tb(ncf, nw, lists, breit, br: "Breit", iab, kstart, keys, vals, lfordr=False, iccut=0):
    """SETHAM with the transverse Breit terms (breit: bt_beg, bt_end, blabel,
    bcoeff; br: the Breit state, advanced as the reference's modules are).
    Returns (row_end, col, val, eav_sum, elsto)."""
    L = lib()
    npairs = len(lists["pair_ic"])
    cap = npairs + 1
    row_end = np.zeros(ncf + 1, dtype=np.int64)
    col = np.zeros(cap, dtype=np.int32)
    val = np.zeros(cap)
    eav = C.c_double(0.0)
    elsto = C.c_double(0.0)
    b = {k: np.ascontiguousarray(breit[k]) for k in ("bt_beg", "bt_end", "blabel", "bcoeff")}
    L.orc_setham_tb.restype = C.c_int64
    nnz = L.orc_setham_tb(C.c_int(ncf), C.c_int(nw), C.c_int64(npairs),
                          i32p(lists["pair_ic"]), i32p(lists["pair_ir"]), i32p(lists["one_ia"]),
                          i32p(lists["one_ib"]), i64p(lists["tsh_off"]), dp(lists["tshell"]),
                          i64p(lists["tb_beg"]), i64p(lists["tb_end"]), i32p(lists["label"]),
                          dp(lists["coeff"]), dp(np.ascontiguousarray(iab, dtype=np.float64)),
                          i32p(kstart), i32p(keys), dp(vals), i64p(b["bt_beg"]), i64p(b["bt_end"]),
                          i32p(b["blabel"]), dp(b["bcoeff"]), C.c_void_p(br.h), C.c_int(int(lfordr)),
                          C.c_int(iccut), i64p(row_end), i32p(col), dp(val), C.byref(eav), C.byref(elsto))
    if nnz < 0:
        raise ValueError("setham_tb: malformed list or missing integral")
    return row_end, col[:nnz].copy(), val[:nnz].copy(), eav.value, elsto.value


def spicmv2(row_end, col, val, B, nthreads: int = 1):
    B = np.asfortranarray(B, dtype=np.float64)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n, m = B.shape
    Cm = np.zeros((n, m), order="F")
    lib().orc_spicmv2_mt(C.c_int(n), C.c_int(m), i64p(row_end), i32p(col), dp(val), dp(B), dp(Cm),
                         C.c_int(nthreads))
    return Cm


def syev(a):
    a = np.array(a, dtype=np.float64, order="F")
    n = a.shape[0]
    w = np.zeros(n)
    rc = lib().orc_syev(C.c_int(n), dp(a), dp(w))
    if rc:
        raise RuntimeError("tql2 did not converge")
    return w, a


def spevx_lowest(a, nev, nthreads: int = 1):
    """Lowest nev eigenpairs of a dense symmetric matrix by DSPEVX's steps
    (tridiagonal reduction, bisection, inverse iteration, back-transform)."""
    a = np.array(a, dtype=np.float64, order="F")
    n = a.shape[0]
    w = np.zeros(nev)
    z = np.zeros((n, nev), order="F")
    rc = lib().orc_spevx_lowest(C.c_int(n), dp(a), C.c_int(nev), dp(w), dp(z), C.c_int(nthreads))
    if rc:
        raise RuntimeError("spevx_lowest failed")
    return w, z


def maneig(row_end, col, val, eav, nvec, nthreads: int = 1):
    """MANEIG for one block: (eval of H - EAV, evec, nloops, nmv, rc)."""
    ncf = len(row_end) - 1
    ev = np.zeros(nvec)
    evec = np.zeros(ncf * nvec)
    nl, nm = C.c_int(0), C.c_int(0)
    rc = lib().orc_maneig(C.c_int(ncf), i64p(row_end), i32p(col), dp(val), C.c_double(eav), C.c_int(nvec),
                          dp(ev), dp(evec), C.c_int(nthreads), C.byref(nl), C.byref(nm))
    return ev, evec.reshape(nvec, ncf).T, nl.value, nm.value, rc


class Breit:
    """rci90's transverse Breit radial integrals (breit_oracle.c): BESSEL's
    cache, the GENINTBREIT1/2 tables and the BRINT3..6 memo, all as
    state-dependent as the reference.  e: orbital energies E (orb_C),
    c: speed of light, wfact: WFACT (photon frequency scale)."""

    def __init__(self, g: "Grid", o: "Orbitals", e, c: float, wfact: float = 1.0):
        L = lib()
        L.orc_breit_new.restype = C.c_void_p
        L.orc_brintf.restype = C.c_double
        L.orc_genintbreit.restype = C.c_int64
        self._keep = (g, o, np.ascontiguousarray(e, dtype=np.float64))
        self.e = self._keep[2]
        self.nw = o.nw
        self.h = L.orc_breit_new(C.c_int(g.n), dp(g.r), dp(g.rp), dp(g.rpor), dp(g.ncc), dp(g.cnc5c),
                                 C.c_double(g.accy), C.c_int(o.nw), C.c_int(o.ld), dp(o.P), dp(o.Q),
                                 i32p(o.mf), i32p(o.nak), i32p(o.nkl), i32p(o.nkj), dp(self.e),
                                 C.c_double(c), C.c_double(wfact))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_breit_free(C.c_void_p(self.h))
            self.h = None

    def brintf(self, itype, ia, ib, ic, id_, k) -> float:
        return lib().orc_brintf(C.c_void_p(self.h), C.c_int(itype), C.c_int(ia), C.c_int(ib), C.c_int(ic),
                                C.c_int(id_), C.c_int(k))

    def genintbreit(self, itype):
        """(kstart[0..j2max+1], keys, vals) of GENINTBREIT1 (1) or 2 (2)."""
        L = lib()
        j2 = C.c_int(0)
        ks = np.zeros(24, dtype=np.int32)
        # count sweep on a scratch copy would change no state: the count
        # needs no BRINTF call, the fill sweep does
        n = L.orc_genintbreit(C.c_void_p(self.h), C.c_int(itype), C.byref(j2), i32p(ks), None, None)
        keys = np.zeros(max(n, 1), dtype=np.int32)
        vals = np.zeros(max(n, 1))
        # the call above filled the table (and advanced BESSEL's cache) once;
        # read it back without recomputing
        L.orc_breit_table.restype = C.c_int64
        L.orc_breit_table(C.c_void_p(self.h), C.c_int(itype), i32p(keys), dp(vals))
        return ks[: j2.value + 2].copy(), keys[:n], vals[:n]

    def brint(self, itype, ia, ib, ic, id_, k) -> float:
        v = C.c_double(0.0)
        if lib().orc_brint(C.c_void_p(self.h), C.c_int(itype), C.c_int(ia), C.c_int(ib), C.c_int(ic),
                           C.c_int(id_), C.c_int(k), C.byref(v)):
            raise KeyError("Breit integral not found")
        return v.value


def maneig_mcdhf(row_end, col, val, eav, nvect, iccmin, warm=None, nthreads: int = 1):
    """rmcdhf90's MANEIG for levels iccmin (1-based).  warm: None (cold
    start) or (eval, evec) of the previous iteration for those levels.
    Returns (eval of H - EAV, evec, nloops, nmv, rc)."""
    ncf = len(row_end) - 1
    lev = np.ascontiguousarray(iccmin, dtype=np.int32)
    nlev = len(lev)
    ev = np.zeros(nlev)
    evec = np.zeros(ncf * nlev)
    if warm is not None:
        ev[:] = warm[0]
        evec[:] = np.asfortranarray(warm[1]).ravel(order="F")
    nl, nm = C.c_int(0), C.c_int(0)
    rc = lib().orc_maneig_mcdhf(C.c_int(ncf), i64p(row_end), i32p(col), dp(val), C.c_double(eav),
                                C.c_int(nvect), C.c_int(nlev), i32p(lev), C.c_int(int(warm is not None)),
                                dp(ev), dp(evec), C.c_int(nthreads), C.byref(nl), C.byref(nm))
    return ev, evec.reshape(nlev, ncf).T, nl.value, nm.value, rc


def gdvd(row_end, col, val, diag, nvec, lim, niv_vectors, mblock=1, crite=1e-17, critc=1e-9,
         critr=1e-9, maxiter=None, nthreads: int = 1):
    """GDVD/DVDSON on a (pre-shifted) CSR; returns (eval, evec, nloops, nmv, ierr)."""
    ncf = len(row_end) - 1
    if maxiter is None:
        maxiter = max(nvec * 200, ncf // 10)
    work = np.zeros(lib().orc_gdvd_worksize(ncf, lim, nvec))
    niv = niv_vectors.shape[1]
    work[: ncf * niv] = np.asfortranarray(niv_vectors).ravel(order="F")
    d = np.array(diag, dtype=np.float64)
    nl, nm = C.c_int(0), C.c_int(0)
    ierr = lib().orc_gdvd(C.c_int(ncf), i64p(row_end), i32p(col), dp(val), C.c_int(lim), dp(d), C.c_int(nvec),
                          C.c_int(niv), C.c_int(mblock), C.c_double(crite), C.c_double(critc),
                          C.c_double(critr), C.c_int(maxiter), dp(work), C.c_int(nthreads), C.byref(nl),
                          C.byref(nm))
    ev = work[ncf * nvec: ncf * nvec + nvec].copy()
    vec = work[: ncf * nvec].reshape(nvec, ncf).T.copy()
    res = work[(ncf + 2) * nvec: (ncf + 2) * nvec + nvec].copy()
    return ev, vec, nl.value, nm.value, ierr, res


def iniest2(row_end, col, val, niv, nmax=2000):
    ncf = len(row_end) - 1
    basis = np.zeros(ncf * niv)
    rc = lib().orc_iniest2(C.c_int(nmax), C.c_int(ncf), C.c_int(niv), i64p(row_end), i32p(col), dp(val),
                           dp(basis))
    if rc:
        raise RuntimeError("iniest2 eigensolve failed")
    return basis.reshape(niv, ncf).T.copy()
