"""GRASP result files (grasp_amd.grasp_io): the .cm mixing file and .res
block rows, against the reference-held He file
src/appl/rhfszeeman95/examples/data_He/3Pn6.cm (copied as the fixture
tests/golden/he_3pn6.cm) and through round trips."""
import os
import struct

import numpy as np
import pytest

from grasp_amd import grasp_io as G
from grasp_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
CM = os.path.join(HERE, "golden", "he_3pn6.cm")
# the same levels as the reference's 3Pn6.csum (EAV + EVAL)
CSUM = {1: [-2.13325863106008], 2: [-2.13326313263103, -2.12392844472579], 3: [-2.13326348364110]}


def test_reference_cm_reads_and_rewrites_byte_for_byte(tmp_path):
    cm = G.read_cm(CM)
    assert (cm.nelec, cm.ncftot, cm.nw, cm.ncmin, cm.nvecsiz) == (2, 472, 36, 4, 652)
    assert [(b.jblock, b.ncf, b.nvec, b.iatjpo, b.iaspar) for b in cm.blocks] == \
        [(1, 70, 1, 1, -1), (2, 180, 2, 3, -1), (3, 222, 1, 5, -1)]
    for b in cm.blocks:
        assert np.allclose(b.energies, CSUM[b.jblock], rtol=0, atol=5e-14)
        assert np.allclose(np.linalg.norm(b.evec, axis=0), 1.0, atol=1e-10)
    out = tmp_path / "copy.cm"
    G.write_cm(out, cm)
    assert open(out, "rb").read() == open(CM, "rb").read()


def test_cm_rejects_truncated_and_foreign_files(tmp_path):
    data = open(CM, "rb").read()
    bad = tmp_path / "bad.cm"
    bad.write_bytes(data[:-3])
    with pytest.raises(ValueError):
        G.read_cm(bad)
    bad.write_bytes(struct.pack("<i", 6) + b"G92RES" + struct.pack("<i", 6))
    with pytest.raises(ValueError):
        G.read_cm(bad)


def test_res_rows_layout_and_round_trip(tmp_path):
    # a 3x3 block: row 1 = (d1), row 2 = (h21, d2), row 3 = (h31, d3)
    iendc = np.array([0, 1, 3, 5])
    irow = np.array([1, 1, 2, 1, 3], dtype=np.int32)
    emt = np.array([-1.0, 0.25, -2.0, 0.5, -3.0])
    p = tmp_path / "b.res"
    G.write_res_block(p, iendc, irow, emt, elsto=0.125, iccut=2)
    raw = open(p, "rb").read()
    # (NCF, ICCUT, MYID, NPROCS) then NELC, ELSTO, EMT(1:NELC), IROW(1:NELC) per row
    want = struct.pack("<i4ii", 16, 3, 2, 0, 1, 16)
    for s, e in ((0, 1), (1, 3), (3, 5)):
        n = e - s
        body = struct.pack("<id", n, 0.125) + emt[s:e].tobytes() + irow[s:e].tobytes()
        want += struct.pack("<i", len(body)) + body + struct.pack("<i", len(body))
    assert raw == want
    blk, = G.read_res_blocks(p)
    assert blk["ncf"] == 3 and blk["iccut"] == 2 and blk["elsto"] == 0.125
    assert np.array_equal(blk["iendc"], iendc) and np.array_equal(blk["irow"], irow)
    assert np.array_equal(blk["emt"], emt)


def test_res_rows_of_several_blocks(tmp_path):
    p = tmp_path / "m.res"
    blocks = [W.clustered_block(n, s, fill=0.1) for n, s in ((50, 1), (120, 2))]
    for i, csr in enumerate(blocks):
        G.write_res_block(p, csr.row_end, csr.col, csr.val, append=i > 0)
    got = G.read_res_blocks(p)
    assert len(got) == 2
    for g, csr in zip(got, blocks):
        assert np.array_equal(g["iendc"], csr.row_end) and np.array_equal(g["irow"], csr.col)
        assert np.array_equal(g["emt"], csr.val)


def test_oracle_levels_written_as_cm_match_reference_file(tmp_path):
    """The oracle's CI on the reference's own angular lists (He+ hydrogenic
    orbitals), written as a .cm file beside the reference's: same layout and
    symmetries, levels within 5 mhartree of the reference's."""
    import oracle as O
    from test_angular_lists import he_tables, lists, G as LISTS

    T = he_tables()
    ref = G.read_cm(CM)
    cm = G.CmFile(nelec=2, nw=36)
    for rb in ref.blocks:
        ncf = int(LISTS[f"b{rb.jblock}_ncf"])
        row_end, col, val, eav_sum = O.setham(ncf, 36, lists(rb.jblock), T["iab"], T["kst"], T["keys"], T["vals"])
        eav = eav_sum / ncf
        w, v, *_ = O.maneig(row_end, col, val, eav, rb.nvec)
        cm.blocks.append(G.CmBlock(rb.jblock, ncf, rb.iatjpo, rb.iaspar, np.arange(1, rb.nvec + 1, dtype=np.int32),
                                   eav, w, v))
    out = tmp_path / "ours.cm"
    G.write_cm(out, cm)
    back = G.read_cm(out)
    assert (back.nelec, back.ncftot, back.nw, back.ncmin, back.nvecsiz) == \
        (ref.nelec, ref.ncftot, ref.nw, ref.ncmin, ref.nvecsiz)
    for b, r in zip(back.blocks, ref.blocks):
        assert (b.jblock, b.ncf, b.nvec, b.iatjpo, b.iaspar) == (r.jblock, r.ncf, r.nvec, r.iatjpo, r.iaspar)
        assert np.all(np.abs(b.energies - r.energies) < 5e-3)
        assert np.allclose(b.evec.T @ b.evec, np.eye(b.nvec), atol=1e-10)
