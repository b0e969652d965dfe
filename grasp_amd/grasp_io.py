"""GRASP2018 result files of the rci90 path, read and written byte for byte.

Both files are Fortran sequential unformatted files: every WRITE is one
record framed by its byte length as a 4-byte integer before and after (the
framing of gfortran and flang on little-endian hosts).

* ``.cm`` -- the mixing-coefficient file (G92MIX), as rci90 writes it:
  SETMIX writes 'G92MIX' (rci90/setmix.f90:48); LODMIX the header
  NELEC, NCFTOT, NW, NCMIN, NVECSIZ, NBLOCK (rci90/lodmix.f90:75); MATRIX,
  per J/parity block, four records (rci90/matrix.f90:209-212):
  (JBLOCK, NCF, NVEC, IATJPO = 2J+1, IASPAR = parity sign), IVEC(1:NVEC),
  (EAV, EVAL(1:NVEC)) and EVEC(NCF*NVEC) column-major.  EVAL is relative to
  EAV, as the file holds it; downstream programs (jj2lsj90, rlevels,
  rtransition90) read it in this layout.
* ``.res`` block rows -- the restart rows SETHAM writes for one block: a
  block record (NCF, ICCUT, MYID, NPROCS) (rci90/genmat.f90:92), then one
  record per row IC: NELC, ELSTO, EMT(1:NELC), IROW(1:NELC), lower-triangle
  columns ascending with the diagonal last (rci90/setham_gg.f90:376-379).
  These rows are exactly what gci_hmat_upload ingests.  The file header that
  SETRES writes before the blocks (grid, orbitals) is not handled here.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

_MARK = "<i"


def _records(data: bytes):
    off = 0
    while off < len(data):
        (n,) = struct.unpack_from(_MARK, data, off)
        if n < 0 or off + 8 + n > len(data):
            raise ValueError(f"malformed record at byte {off}")
        body = data[off + 4: off + 4 + n]
        (n2,) = struct.unpack_from(_MARK, data, off + 4 + n)
        if n2 != n:
            raise ValueError(f"record markers disagree at byte {off}")
        yield body
        off += 8 + n


def _record(body: bytes) -> bytes:
    return struct.pack(_MARK, len(body)) + body + struct.pack(_MARK, len(body))


@dataclass
class CmBlock:
    jblock: int
    ncf: int
    iatjpo: int          # 2J + 1
    iaspar: int          # parity: +1 even, -1 odd
    ivec: np.ndarray     # NVEC level indices (1-based)
    eav: float
    eval: np.ndarray     # NVEC eigenvalues relative to EAV
    evec: np.ndarray     # NCF x NVEC, column-major

    @property
    def nvec(self) -> int:
        return len(self.ivec)

    @property
    def energies(self) -> np.ndarray:
        """Total level energies EAV + EVAL (rci90/engout.f90)."""
        return self.eav + self.eval


@dataclass
class CmFile:
    nelec: int
    nw: int
    blocks: list = field(default_factory=list)

    @property
    def ncftot(self) -> int:
        return sum(b.ncf for b in self.blocks)

    @property
    def ncmin(self) -> int:
        return sum(b.nvec for b in self.blocks)

    @property
    def nvecsiz(self) -> int:
        return sum(b.ncf * b.nvec for b in self.blocks)


def read_cm(path) -> CmFile:
    recs = list(_records(open(path, "rb").read()))
    if not recs or recs[0] != b"G92MIX":
        raise ValueError("not a G92MIX file")
    nelec, ncftot, nw, ncmin, nvecsiz, nblock = struct.unpack("<6i", recs[1])
    cm = CmFile(nelec, nw)
    r = 2
    for _ in range(nblock):
        jb, ncf, nvec, iatjpo, iaspar = struct.unpack("<5i", recs[r])
        ivec = np.frombuffer(recs[r + 1], dtype="<i4").copy()
        ev = np.frombuffer(recs[r + 2], dtype="<f8")
        evec = np.frombuffer(recs[r + 3], dtype="<f8").reshape(nvec, ncf).T.copy()
        if len(ivec) != nvec or len(ev) != nvec + 1:
            raise ValueError(f"block {jb}: record sizes do not match NVEC = {nvec}")
        cm.blocks.append(CmBlock(jb, ncf, iatjpo, iaspar, ivec, float(ev[0]), ev[1:].copy(), evec))
        r += 4
    if (cm.ncftot, cm.ncmin, cm.nvecsiz) != (ncftot, ncmin, nvecsiz):
        raise ValueError("header counts do not match the blocks")
    return cm


def write_cm(path, cm: CmFile):
    out = [_record(b"G92MIX"),
           _record(struct.pack("<6i", cm.nelec, cm.ncftot, cm.nw, cm.ncmin, cm.nvecsiz, len(cm.blocks)))]
    for b in cm.blocks:
        out.append(_record(struct.pack("<5i", b.jblock, b.ncf, b.nvec, b.iatjpo, b.iaspar)))
        out.append(_record(np.ascontiguousarray(b.ivec, dtype="<i4").tobytes()))
        out.append(_record(np.concatenate([[b.eav], np.asarray(b.eval, dtype=np.float64)]).astype("<f8").tobytes()))
        out.append(_record(np.asarray(b.evec, dtype="<f8").T.tobytes()))  # column-major
    with open(path, "wb") as f:
        f.write(b"".join(out))


def write_res_block(path, iendc, irow, emt, elsto: float = 0.0, iccut: int = 0, myid: int = 0, nprocs: int = 1,
                    append: bool = False):
    """One block's restart rows: (NCF, ICCUT, MYID, NPROCS), then per row
    NELC, ELSTO, EMT(1:NELC), IROW(1:NELC) (lower rows, diagonal last)."""
    iendc = np.asarray(iendc, dtype=np.int64)
    irow = np.asarray(irow, dtype="<i4")
    emt = np.asarray(emt, dtype="<f8")
    ncf = len(iendc) - 1
    out = [_record(struct.pack("<4i", ncf, iccut, myid, nprocs))]
    for i in range(ncf):
        s, e = int(iendc[i]), int(iendc[i + 1])
        out.append(_record(struct.pack("<id", e - s, elsto) + emt[s:e].tobytes() + irow[s:e].tobytes()))
    with open(path, "ab" if append else "wb") as f:
        f.write(b"".join(out))


def read_res_blocks(path):
    """Blocks of restart rows written by write_res_block (or SETHAM): a list
    of dicts with ncf, iccut, myid, nprocs, elsto and the rows as
    (iendc, irow, emt) in the layout gci_hmat_upload takes."""
    recs = list(_records(open(path, "rb").read()))
    blocks, r = [], 0
    while r < len(recs):
        ncf, iccut, myid, nprocs = struct.unpack("<4i", recs[r])
        r += 1
        iendc = np.zeros(ncf + 1, dtype=np.int64)
        cols, vals, elsto = [], [], 0.0
        for i in range(ncf):
            body = recs[r + i]
            nelc, elsto = struct.unpack_from("<id", body, 0)
            if len(body) != 12 + 12 * nelc:
                raise ValueError(f"row {i + 1}: record length {len(body)} does not match NELC = {nelc}")
            vals.append(np.frombuffer(body, dtype="<f8", count=nelc, offset=12))
            cols.append(np.frombuffer(body, dtype="<i4", count=nelc, offset=12 + 8 * nelc))
            iendc[i + 1] = iendc[i] + nelc
        r += ncf
        blocks.append({"ncf": ncf, "iccut": iccut, "myid": myid, "nprocs": nprocs, "elsto": elsto,
                       "iendc": iendc, "irow": np.concatenate(cols) if cols else np.zeros(0, np.int32),
                       "emt": np.concatenate(vals) if vals else np.zeros(0)})
    return blocks
