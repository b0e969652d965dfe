"""Host-side mirror of the rci90 hot-path interface, over libgraspci.

Names and argument meanings follow the GRASP2018 routines each call
replaces; errors raise ``GciError`` where the Fortran would STOP.

* ``HamiltonianBlock`` -- the hmat_C module state (EMT, IROW, IENDC,
  lib/libmod/hmat_C.f90:6-9) as a device-resident handle, loaded the way
  MANEIG loads it (rci90/maneig.f90:312-332: diagonal shifted by -EAV).
* ``spicmv2`` -- the GDVD operator OP(N, M, B, C) (lib9290/spicmv2.f90:2).
* ``gdvd`` / ``maneig`` -- Davidson and the MANEIG dispatch.
* ``genintrk`` / ``one_body`` / ``setham`` -- R^k table, I(ab), assembly.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, dp, i32p, i64p


class HamiltonianBlock:
    """Device-resident block Hamiltonian (lower rows, diagonal last)."""

    def __init__(self, handle, n: int, nnz: int, precision: int):
        self._h = handle
        self.n = n
        self.nnz_offdiag = nnz
        self.precision = precision

    @classmethod
    def from_rows(cls, iendc, irow, emt, shift: float = 0.0, precision: int = L.GCI_FP64,
                  index_base: int = 1):
        iendc = np.ascontiguousarray(iendc, dtype=np.int64)
        irow = np.ascontiguousarray(irow, dtype=np.int32)
        emt = np.ascontiguousarray(emt, dtype=np.float64)
        n = len(iendc) - 1
        h = C.c_void_p()
        check(L.lib().gci_hmat_upload(C.c_int(n), i64p(iendc), i32p(irow), dp(emt), C.c_int(index_base),
                                      C.c_double(shift), C.c_int(precision), C.byref(h)))
        nn, nz, pr = C.c_int64(), C.c_int64(), C.c_int()
        check(L.lib().gci_hmat_info(h, C.byref(nn), C.byref(nz), C.byref(pr)))
        return cls(h, nn.value, nz.value, pr.value)

    @classmethod
    def from_csr(cls, csr, shift: float = 0.0, precision: int = L.GCI_FP64):
        """From a workloads.LowerCSR (1-based columns, diagonal last)."""
        return cls.from_rows(csr.row_end, csr.col, csr.val, shift, precision, 1)

    @property
    def handle(self):
        return self._h

    def free(self):
        if self._h:
            check(L.lib().gci_hmat_free(self._h))
            self._h = None

    def release(self):
        """Give up ownership (the server frees it)."""
        h, self._h = self._h, None
        return h

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def download(self, index_base: int = 1):
        n = self.n
        iendc = np.zeros(n + 1, dtype=np.int64)
        irow = np.zeros(self.nnz_offdiag + n, dtype=np.int32)
        emt = np.zeros(self.nnz_offdiag + n)
        check(L.lib().gci_hmat_download(self._h, i64p(iendc), i32p(irow), dp(emt), C.c_int(index_base)))
        return iendc, irow, emt

    def spicmv2(self, B, mode: int = L.GCI_OP_FP64):
        """C := H B (SPICMV2 semantics), B column-major n x m (or a vector)."""
        B = np.asarray(B, dtype=np.float64)
        vec = B.ndim == 1
        Bf = np.asfortranarray(B.reshape(self.n, -1) if vec else B)
        m = Bf.shape[1]
        Cf = np.zeros((self.n, m), order="F")
        check(L.lib().gci_spmm(self._h, C.c_int(m), dp(Bf), dp(Cf), C.c_int(mode)))
        return Cf[:, 0] if vec else Cf


def spicmv2(block: HamiltonianBlock, n: int, m: int, B, C_out=None, mode: int = L.GCI_OP_FP64):
    """SPICMV2(N, M, B, C) with the block standing in for module hmat_C."""
    if n != block.n:
        raise ValueError("spicmv2: N does not match the loaded matrix")
    res = block.spicmv2(np.asarray(B).reshape(n, m, order="F"), mode)
    if C_out is not None:
        C_out[...] = res.reshape(C_out.shape, order="F")
        return C_out
    return res


def gdvd(block: HamiltonianBlock, nvec: int, lim: int = 0, mblock: int = 1, crite: float = 1.0e-17,
         critc: float = 1.0e-9, critr: float = 1.0e-9, maxiter: int = 0, init=None,
         mode: int = L.GCI_OP_FP64, refine: bool = True):
    """GDVD/DVDSON (libdvd90/gdvd.f90) for the lowest nvec eigenpairs of the
    block as held.  init: n x niv initial estimates (None = INIEST2).
    Returns (eval, evec, resid, stats dict)."""
    n = block.n
    evec = np.zeros((n, nvec), order="F")
    niv = 0
    if init is not None:
        init = np.asfortranarray(init, dtype=np.float64)
        niv = init.shape[1]
        if niv > nvec:
            raise ValueError("gdvd: more initial vectors than nvec")
        evec[:, :niv] = init
    p = L.GciDvdParams(nvec, lim, mblock, maxiter, crite, critc, critr, niv, mode, int(refine))
    ev = np.zeros(nvec)
    res = np.zeros(nvec)
    st = L.GciDvdStats()
    check(L.lib().gci_davidson(block.handle, C.byref(p), dp(ev), dp(evec), dp(res), C.byref(st)))
    return ev, evec, res, {"nloops": st.nloops, "nmv": st.nmv, "ierr": st.ierr,
                           "refine_loops": st.refine_loops, "seconds": st.seconds}


def maneig(block: HamiltonianBlock, nvec: int):
    """MANEIG for one block (rci90/maneig.f90): dense for NCF <= 2000, else
    INIEST2 + GDVD(SPICMV2); eigenvectors sign-fixed."""
    n = block.n
    ev = np.zeros(nvec)
    evec = np.zeros((n, nvec), order="F")
    st = L.GciDvdStats()
    check(L.lib().gci_maneig(block.handle, C.c_int(nvec), dp(ev), dp(evec), C.byref(st)))
    return ev, evec, {"nloops": st.nloops, "nmv": st.nmv, "ierr": st.ierr, "seconds": st.seconds}


def maneig_mcdhf(block: HamiltonianBlock, nvect: int, iccmin, warm=None):
    """MANEIG of the SCF program (rmcdhf90/maneig.f90:78-200): the levels
    iccmin (1-based) of the lowest min(nvect, NCF); warm = (eval, evec) of
    the previous SCF iteration for those levels, or None for a cold start
    (INIEST2 over the leading 4000 CSFs).  Returns (eval, evec, stats)."""
    n = block.n
    lev = np.ascontiguousarray(iccmin, dtype=np.int32)
    nlev = len(lev)
    ev = np.zeros(nlev)
    evec = np.zeros((n, nlev), order="F")
    if warm is not None:
        ev[:] = warm[0]
        evec[:] = warm[1]
    st = L.GciDvdStats()
    check(L.lib().gci_maneig_mcdhf(block.handle, C.c_int(nvect), C.c_int(nlev), i32p(lev),
                                   C.c_int(int(warm is not None)), dp(ev), dp(evec), C.byref(st)))
    return ev, evec, {"nloops": st.nloops, "nmv": st.nmv, "ierr": st.ierr, "seconds": st.seconds}


def _grid_struct(r, rp, rpor, n, h):
    r = np.ascontiguousarray(r, dtype=np.float64)
    rp = np.ascontiguousarray(rp, dtype=np.float64)
    rpor = np.ascontiguousarray(rpor, dtype=np.float64)
    g = L.GciGrid(n, h, h ** 6, dp(r), dp(rp), dp(rpor))
    return g, (r, rp, rpor)


def _orb_struct(P, Q, mf, nak, nkl, nkj):
    arrs = (np.ascontiguousarray(P, dtype=np.float64), np.ascontiguousarray(Q, dtype=np.float64),
            np.ascontiguousarray(mf, dtype=np.int32), np.ascontiguousarray(nak, dtype=np.int32),
            np.ascontiguousarray(nkl, dtype=np.int32), np.ascontiguousarray(nkj, dtype=np.int32))
    nw, ld = arrs[0].shape
    o = L.GciOrbitals(nw, ld, dp(arrs[0]), dp(arrs[1]), i32p(arrs[2]), i32p(arrs[3]), i32p(arrs[4]),
                      i32p(arrs[5]))
    return o, arrs


def radgrd(n: int = 590, rnt: float = 2.0e-6, h: float = 0.05):
    """RADGRD (lib9290/radgrd.f90:41-60): r, rp, rpor on n+10 points."""
    r, rp, rpor = np.zeros(n + 10), np.zeros(n + 10), np.zeros(n + 10)
    check(L.lib().gci_radgrd(C.c_int(n), C.c_double(rnt), C.c_double(h), dp(r), dp(rp), dp(rpor)))
    return r, rp, rpor


def genintrk(r, rp, rpor, n, h, orbitals):
    """GENINTRK (rci90/genintrk.f90): returns (kstart, keys, vals) with the
    reference's enumeration, packing and KSTART; values computed on the GPU."""
    g, keep_g = _grid_struct(r, rp, rpor, n, h)
    o, keep_o = _orb_struct(orbitals.P, orbitals.Q, orbitals.mf, orbitals.nak, orbitals.nkl, orbitals.nkj)
    j2 = C.c_int32(0)
    cnt = C.c_int64(0)
    kst = np.zeros(22, dtype=np.int32)
    check(L.lib().gci_rk_table(C.byref(g), C.byref(o), C.byref(j2), i32p(kst), None, None, C.c_int64(0),
                               C.byref(cnt)))
    nk = cnt.value
    keys = np.zeros(max(nk, 1), dtype=np.int32)
    vals = np.zeros(max(nk, 1))
    check(L.lib().gci_rk_table(C.byref(g), C.byref(o), C.byref(j2), i32p(kst), i32p(keys), dp(vals),
                               C.c_int64(nk), C.byref(cnt)))
    return kst[: j2.value + 2].copy(), keys[:nk], vals[:nk]


def yzk(r, rp, rpor, n, h, orbitals, k, b, d):
    """YZK (lib9290/yzk.f90) for a batch of (k, b, d), 1-based orbitals:
    returns (yk[npairs, n], mtp[npairs]), TB(1:N) and MTP per pair."""
    g, keep_g = _grid_struct(r, rp, rpor, n, h)
    o, keep_o = _orb_struct(orbitals.P, orbitals.Q, orbitals.mf, orbitals.nak, orbitals.nkl, orbitals.nkj)
    k = np.ascontiguousarray(k, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    d = np.ascontiguousarray(d, dtype=np.int32)
    npairs = len(k)
    yk = np.zeros((npairs, n))
    mtp = np.zeros(npairs, dtype=np.int32)
    check(L.lib().gci_yzk(C.byref(g), C.byref(o), C.c_int(npairs), i32p(k), i32p(b), i32p(d), dp(yk), i32p(mtp)))
    return yk, mtp


def one_body(r, rp, rpor, n, h, orbitals, zz, c):
    """I(a,b) = RINTI(a,b,0) for same-kappa pairs a <= b (nw x nw, upper)."""
    g, keep_g = _grid_struct(r, rp, rpor, n, h)
    o, keep_o = _orb_struct(orbitals.P, orbitals.Q, orbitals.mf, orbitals.nak, orbitals.nkl, orbitals.nkj)
    nw = keep_o[0].shape[0]
    iab = np.zeros((nw, nw))
    zz = np.ascontiguousarray(zz, dtype=np.float64)
    check(L.lib().gci_one_body(C.byref(g), C.byref(o), dp(zz), C.c_double(c), dp(iab)))
    return iab


class BreitTables:
    """rci90's transverse Breit radial integrals on the GPU
    (gci_breit_create: GENINTBREIT1/2, then BRINT3..6 on demand), with the
    reference's module state carried from block to block.  e: orbital
    energies E, c: speed of light, wfact: WFACT."""

    def __init__(self, r, rp, rpor, n, h, orbitals, e, c, wfact=1.0):
        self._g = _grid_struct(r, rp, rpor, n, h)
        self._o = _orb_struct(orbitals.P, orbitals.Q, orbitals.mf, orbitals.nak, orbitals.nkl, orbitals.nkj)
        self._e = np.ascontiguousarray(e, dtype=np.float64)
        self.h = C.c_void_p()
        check(L.lib().gci_breit_create(C.byref(self._g[0]), C.byref(self._o[0]), dp(self._e), C.c_double(c),
                                       C.c_double(wfact), C.byref(self.h)))

    def table(self, itype):
        """(kstart[0..j2max+1], keys, vals) of GENINTBREIT1 (1) or 2 (2)."""
        j2 = C.c_int32(0)
        cnt = C.c_int64(0)
        kst = np.zeros(22, dtype=np.int32)
        check(L.lib().gci_breit_table(self.h, C.c_int(itype), C.byref(j2), i32p(kst), None, None, C.c_int64(0),
                                      C.byref(cnt)))
        nk = cnt.value
        keys = np.zeros(max(nk, 1), dtype=np.int32)
        vals = np.zeros(max(nk, 1))
        check(L.lib().gci_breit_table(self.h, C.c_int(itype), C.byref(j2), i32p(kst), i32p(keys), dp(vals),
                                      C.c_int64(nk), C.byref(cnt)))
        return kst[: j2.value + 2].copy(), keys[:nk], vals[:nk]

    def integral(self, itype, ia, ib, ic, id_, k) -> float:
        """BRINTn(ia, ib, ic, id, k) as SETHAM calls it."""
        v = C.c_double(0.0)
        check(L.lib().gci_breit_integral(self.h, C.c_int(itype), C.c_int(ia), C.c_int(ib), C.c_int(ic),
                                         C.c_int(id_), C.c_int(k), C.byref(v)))
        return v.value

    def free(self):
        if self.h:
            L.lib().gci_breit_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def setham(ncf, nw, lists, iab, kstart, keys, vals, lfordr=False, iccut=0, precision=L.GCI_FP64, breit=None,
           tables=None):
    """SETHAM element loop on the GPU: returns (HamiltonianBlock with the
    diagonal shifted by -EAV, EAV, NELMNT).  With breit (the lists' bt_beg,
    bt_end, blabel, bcoeff) and tables (BreitTables), the transverse Breit
    terms too; EAV then includes ELSTO, which is returned fourth."""
    arrs = {
        "pair_ic": np.ascontiguousarray(lists["pair_ic"], dtype=np.int32),
        "pair_ir": np.ascontiguousarray(lists["pair_ir"], dtype=np.int32),
        "one_ia": np.ascontiguousarray(lists["one_ia"], dtype=np.int32),
        "one_ib": np.ascontiguousarray(lists["one_ib"], dtype=np.int32),
        "tsh_off": np.ascontiguousarray(lists["tsh_off"], dtype=np.int64),
        "tshell": np.ascontiguousarray(lists["tshell"], dtype=np.float64),
        "tb_beg": np.ascontiguousarray(lists["tb_beg"], dtype=np.int64),
        "tb_end": np.ascontiguousarray(lists["tb_end"], dtype=np.int64),
        "label": np.ascontiguousarray(lists["label"], dtype=np.int32),
        "coeff": np.ascontiguousarray(lists["coeff"], dtype=np.float64),
    }
    pl = L.GciPairLists(len(arrs["pair_ic"]), i32p(arrs["pair_ic"]), i32p(arrs["pair_ir"]), i32p(arrs["one_ia"]),
                        i32p(arrs["one_ib"]), i64p(arrs["tsh_off"]), dp(arrs["tshell"]), len(arrs["tshell"]),
                        i64p(arrs["tb_beg"]), i64p(arrs["tb_end"]), i32p(arrs["label"]), dp(arrs["coeff"]),
                        len(arrs["coeff"]))
    iab = np.ascontiguousarray(iab, dtype=np.float64)
    kstart = np.ascontiguousarray(kstart, dtype=np.int32)
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    h = C.c_void_p()
    eav = C.c_double(0.0)
    nel = C.c_int64(0)
    if breit is None:
        check(L.lib().gci_h_assemble(C.c_int(ncf), C.c_int(nw), C.byref(pl), dp(iab), i32p(kstart), i32p(keys),
                                     dp(vals), C.c_int64(len(keys)), C.c_int(int(lfordr)), C.c_int(iccut),
                                     C.c_int(precision), C.byref(h), C.byref(eav), C.byref(nel)))
    else:
        b = {"bt_beg": np.ascontiguousarray(breit["bt_beg"], dtype=np.int64),
             "bt_end": np.ascontiguousarray(breit["bt_end"], dtype=np.int64),
             "blabel": np.ascontiguousarray(breit["blabel"], dtype=np.int32),
             "bcoeff": np.ascontiguousarray(breit["bcoeff"], dtype=np.float64)}
        bl = L.GciBreitLists(i64p(b["bt_beg"]), i64p(b["bt_end"]), i32p(b["blabel"]), dp(b["bcoeff"]),
                             len(b["bcoeff"]))
        elsto = C.c_double(0.0)
        check(L.lib().gci_h_assemble_tb(C.c_int(ncf), C.c_int(nw), C.byref(pl), C.byref(bl), tables.h, dp(iab),
                                        i32p(kstart), i32p(keys), dp(vals), C.c_int64(len(keys)),
                                        C.c_int(int(lfordr)), C.c_int(iccut), C.c_int(precision), C.byref(h),
                                        C.byref(eav), C.byref(elsto), C.byref(nel)))
    nn, nz, pr = C.c_int64(), C.c_int64(), C.c_int()
    check(L.lib().gci_hmat_info(h, C.byref(nn), C.byref(nz), C.byref(pr)))
    blk = HamiltonianBlock(h, nn.value, nz.value, pr.value)
    if breit is None:
        return blk, eav.value, nel.value
    return blk, eav.value, nel.value, elsto.value


def _pair_struct(lists, lo, hi, t_lo, t_hi, s_lo, s_hi):
    """gci_pair_lists for pairs [lo, hi) whose two-body entries are
    [t_lo, t_hi) and TSHELL values [s_lo, s_hi) (offsets made chunk-local)."""
    a = {
        "pair_ic": np.ascontiguousarray(lists["pair_ic"][lo:hi], dtype=np.int32),
        "pair_ir": np.ascontiguousarray(lists["pair_ir"][lo:hi], dtype=np.int32),
        "one_ia": np.ascontiguousarray(lists["one_ia"][lo:hi], dtype=np.int32),
        "one_ib": np.ascontiguousarray(lists["one_ib"][lo:hi], dtype=np.int32),
        "tsh_off": np.ascontiguousarray(np.asarray(lists["tsh_off"][lo:hi]) - s_lo, dtype=np.int64),
        "tshell": np.ascontiguousarray(lists["tshell"][s_lo:s_hi], dtype=np.float64),
        "tb_beg": np.ascontiguousarray(np.asarray(lists["tb_beg"][lo:hi]) - t_lo, dtype=np.int64),
        "tb_end": np.ascontiguousarray(np.asarray(lists["tb_end"][lo:hi]) - t_lo, dtype=np.int64),
        "label": np.ascontiguousarray(lists["label"][5 * t_lo:5 * t_hi], dtype=np.int32),
        "coeff": np.ascontiguousarray(lists["coeff"][t_lo:t_hi], dtype=np.float64),
    }
    pl = L.GciPairLists(hi - lo, i32p(a["pair_ic"]), i32p(a["pair_ir"]), i32p(a["one_ia"]), i32p(a["one_ib"]),
                        i64p(a["tsh_off"]), dp(a["tshell"]), len(a["tshell"]), i64p(a["tb_beg"]), i64p(a["tb_end"]),
                        i32p(a["label"]), dp(a["coeff"]), len(a["coeff"]))
    return pl, a


def setham_streamed(ncf, nw, lists, iab, kstart, keys, vals, chunk_pairs, lfordr=False, iccut=0,
                    precision=L.GCI_FP64, breit=None, tables=None):
    """SETHAM through gci_hstream_*: the lists pushed in chunks of about
    chunk_pairs pairs.  Returns (HamiltonianBlock, EAV, NELMNT, ELSTO)."""
    iab = np.ascontiguousarray(iab, dtype=np.float64)
    kstart = np.ascontiguousarray(kstart, dtype=np.int32)
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    hs = C.c_void_p()
    check(L.lib().gci_hstream_begin(C.c_int(ncf), C.c_int(nw), dp(iab), i32p(kstart), i32p(keys), dp(vals),
                                    C.c_int64(len(keys)), C.c_int(int(lfordr)), C.c_int(iccut), C.c_int(precision),
                                    tables.h if tables is not None else None, C.byref(hs)))
    try:
        npairs = len(lists["pair_ic"])
        tsh_off = np.asarray(lists["tsh_off"])
        tsh_len = np.where(np.asarray(lists["one_ia"]) == 0, 0,
                           np.where(np.asarray(lists["one_ia"]) == np.asarray(lists["one_ib"]), nw, 1))
        keep = []
        for lo in range(0, npairs, chunk_pairs):
            hi = min(npairs, lo + chunk_pairs)
            t_lo, t_hi = int(lists["tb_beg"][lo]), int(lists["tb_end"][hi - 1])
            s_lo = int(tsh_off[lo])
            s_hi = int(max(s_lo, (tsh_off[lo:hi] + tsh_len[lo:hi]).max()))
            pl, a = _pair_struct(lists, lo, hi, t_lo, t_hi, s_lo, s_hi)
            bptr = None
            if breit is not None:
                b_lo, b_hi = int(breit["bt_beg"][lo]), int(breit["bt_end"][hi - 1])
                b = {"bt_beg": np.ascontiguousarray(np.asarray(breit["bt_beg"][lo:hi]) - b_lo, dtype=np.int64),
                     "bt_end": np.ascontiguousarray(np.asarray(breit["bt_end"][lo:hi]) - b_lo, dtype=np.int64),
                     "blabel": np.ascontiguousarray(breit["blabel"][6 * b_lo:6 * b_hi], dtype=np.int32),
                     "bcoeff": np.ascontiguousarray(breit["bcoeff"][b_lo:b_hi], dtype=np.float64)}
                bl = L.GciBreitLists(i64p(b["bt_beg"]), i64p(b["bt_end"]), i32p(b["blabel"]), dp(b["bcoeff"]),
                                     len(b["bcoeff"]))
                keep.append(b)
                bptr = C.byref(bl)
            check(L.lib().gci_hstream_push(hs, C.byref(pl), bptr))
            keep.append(a)
        h = C.c_void_p()
        eav, elsto, nel = C.c_double(0.0), C.c_double(0.0), C.c_int64(0)
        check(L.lib().gci_hstream_end(hs, C.byref(h), C.byref(eav), C.byref(elsto), C.byref(nel)))
    finally:
        L.lib().gci_hstream_free(hs)
    nn, nz, pr = C.c_int64(), C.c_int64(), C.c_int()
    check(L.lib().gci_hmat_info(h, C.byref(nn), C.byref(nz), C.byref(pr)))
    return HamiltonianBlock(h, nn.value, nz.value, pr.value), eav.value, nel.value, elsto.value


class _PinnedBlock:
    """Owner of a page-locked host allocation (freed with the last view)."""

    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        check(L.lib().gci_host_alloc(C.c_size_t(nbytes), C.byref(self.ptr)))

    def __del__(self):
        if self.ptr:
            L.lib().gci_host_free(self.ptr)
            self.ptr = C.c_void_p()


def pinned_empty(count: int) -> np.ndarray:
    """An uninitialised float64 array of `count` elements in page-locked host
    memory: the device writes results into it by DMA, with no staging copy."""
    owner = _PinnedBlock(max(1, int(count)) * 8)
    buf = (C.c_double * max(1, int(count))).from_address(owner.ptr.value)
    buf._owner = owner  # the memory lives while any view of it does
    return np.frombuffer(buf, dtype=np.float64, count=int(count))


def batch_solve(blocks, nev, tol, eav=None, out=None):
    """Batched form: all blocks' Davidson iterations interleaved on the
    device.  blocks: list of HamiltonianBlock; nev/tol/eav: per block.
    Returns (list of eigenvalue arrays, list of eigenvector arrays, stats);
    the eigenvector arrays are n x nev column-major views into one buffer,
    `out` when given (float64, at least sum(n * nev) elements; a
    `pinned_empty` buffer takes the copy directly)."""
    nb = len(blocks)
    nev = np.ascontiguousarray(np.broadcast_to(np.asarray(nev, dtype=np.int32), (nb,)))
    tol = np.ascontiguousarray(np.broadcast_to(np.asarray(tol, dtype=np.float64), (nb,)))
    eav = np.zeros(nb) if eav is None else np.ascontiguousarray(eav, dtype=np.float64)
    handles = (C.c_void_p * nb)(*[b.handle for b in blocks])
    ev = np.zeros(int(nev.sum()))
    total = int(sum(b.n * int(k) for b, k in zip(blocks, nev)))
    if out is None:
        vec = np.empty(total)
    else:
        if out.dtype != np.float64 or out.ndim != 1 or not out.flags.c_contiguous or out.size < total:
            raise ValueError(f"batch_solve: out must be a contiguous float64 vector of at least {total} elements")
        vec = out[:total]
    eo = np.zeros(nb, dtype=np.int64)
    vo = np.zeros(nb, dtype=np.int64)
    st = (L.GciDvdStats * nb)()
    check(L.lib().gci_batch_solve(C.c_int(nb), handles, i32p(nev), dp(tol), dp(eav), dp(ev), i64p(eo), dp(vec),
                                  i64p(vo), st))
    # plain ints for the per-block slicing (numpy scalars index slowly)
    eol, vol, nevl = eo.tolist(), vo.tolist(), nev.tolist()
    evals = [ev[o: o + k] for o, k in zip(eol, nevl)]  # views of this call's own array
    # n x nev column-major views into the one returned buffer (no copy)
    evecs = [vec[o: o + b.n * k].reshape(k, b.n).T for o, b, k in zip(vol, blocks, nevl)]
    keys = [f[0] for f in L.GciDvdStats._fields_]
    stats = [dict(zip(keys, r)) for r in np.ctypeslib.as_array(st).tolist()]
    return evals, evecs, stats


_SHAPE_KEYS = ("R", "block", "nslice", "depth", "target")


def spmm_shape() -> dict:
    """The SpMM launch shape in force: panel height R, workgroup size
    (block), slices per work item (nslice), steps in flight per wavefront
    (depth) and entries per work item (target)."""
    s = L.GciSpmmShape()
    check(L.lib().gci_spmm_get_shape(C.byref(s)))
    return {k: int(getattr(s, k)) for k in _SHAPE_KEYS}


def set_spmm_shape(**kw) -> dict:
    """Changes the given fields of the launch shape; returns the new shape."""
    cur = spmm_shape()
    unknown = set(kw) - set(_SHAPE_KEYS)
    if unknown:
        raise ValueError(f"unknown shape fields {sorted(unknown)}")
    cur.update({k: int(v) for k, v in kw.items()})
    s = L.GciSpmmShape(**cur)
    check(L.lib().gci_spmm_set_shape(C.byref(s)))
    return cur


def spmm_layout(block) -> dict:
    """The block's panel-slice layout for the shape in force: column
    segments, slices, stored entries (padding included), the laid-out
    entries (both triangles for a split-precision block, the lower one
    otherwise) and the padding ratio stored / laid out."""
    ns, nsl, st = C.c_int64(), C.c_int64(), C.c_int64()
    check(L.lib().gci_spmm_layout(block.handle, C.byref(ns), C.byref(nsl), C.byref(st)))
    laid = block.nnz_offdiag * (2 if block.precision == L.GCI_SPLIT32 else 1)
    return {"segments": ns.value, "slices": nsl.value, "stored": st.value, "laid_out": laid,
            "overhead": st.value / max(1, laid)}


def spmm_time(blocks, m: int, mode: int = L.GCI_OP_SPLIT_BULK, reps: int = 5) -> float:
    """Mean ms of one m-column operator application over all blocks."""
    handles = (C.c_void_p * len(blocks))(*[b.handle for b in blocks])
    ms = C.c_double()
    check(L.lib().gci_spmm_time(handles, C.c_int(len(blocks)), C.c_int(m), C.c_int(mode), C.c_int(reps),
                                C.byref(ms)))
    return ms.value


def autotune(blocks, m: int, mode: int = L.GCI_OP_SPLIT_BULK, reps: int = 3):
    """Coordinate search of the launch shape on these blocks; leaves the
    best shape in force.  Returns (shape dict, ms per pass)."""
    handles = (C.c_void_p * len(blocks))(*[b.handle for b in blocks])
    s = L.GciSpmmShape()
    ms = C.c_double()
    check(L.lib().gci_spmm_autotune(handles, C.c_int(len(blocks)), C.c_int(m), C.c_int(mode), C.c_int(reps),
                                    C.byref(s), C.byref(ms)))
    return {k: int(getattr(s, k)) for k in _SHAPE_KEYS}, ms.value


class Server:
    """Continuous-batching CI solve server (gci_server_*)."""

    def __init__(self, hbm_budget_bytes: int, max_active: int = 64):
        self._s = C.c_void_p()
        check(L.lib().gci_server_create(C.c_size_t(int(hbm_budget_bytes)), C.c_int(max_active),
                                        C.byref(self._s)))
        self._meta = {}

    def submit(self, block: HamiltonianBlock, nev: int, tol: float, eav: float = 0.0, shared: bool = False) -> int:
        """Queues one solve.  The server takes the block over, unless `shared`:
        then the caller keeps it (alive until the request is done) and may
        submit it again."""
        req = C.c_int64(0)
        fn = L.lib().gci_server_submit_shared if shared else L.lib().gci_server_submit
        check(fn(self._s, block.handle, C.c_int(nev), C.c_double(tol), C.c_double(eav), C.byref(req)))
        self._meta[req.value] = (block.n, nev)
        if not shared:
            block.release()  # the server owns the matrix now
        return req.value

    def poll(self, req: int):
        """(eval, evec, latency_s) when done, None while running."""
        n, nev = self._meta[req]
        ev = np.zeros(nev)
        vec = np.zeros(n * nev)
        lat = C.c_double(0.0)
        code = L.lib().gci_server_poll(self._s, C.c_int64(req), dp(ev), dp(vec), C.byref(lat))
        if code == L.GCI_ERR_BUSY:
            return None
        check(code)
        del self._meta[req]
        return ev, vec.reshape(nev, n).T.copy(), lat.value

    def drain(self):
        check(L.lib().gci_server_drain(self._s))

    def close(self):
        if self._s:
            check(L.lib().gci_server_destroy(self._s))
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
