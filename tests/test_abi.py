"""The C-ABI library loads and exports every entry point include/graspci.h
declares (no device calls)."""
import ctypes as C
import os
import re

import pytest

from grasp_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "graspci.h")).read()
    return sorted(set(re.findall(r"\b(gci_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_python_export_list():
    assert _declared() == sorted(L.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    for name in _declared():
        assert hasattr(lib, name), name
        assert isinstance(getattr(lib, name), C._CFuncPtr)


def test_library_has_gfx950_code_object():
    data = open(L.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_errors_are_reported_without_a_device():
    # argument validation happens before any device call
    lib = L.lib()
    code = lib.gci_radgrd(C.c_int(0), C.c_double(1.0), C.c_double(0.05), None, None, None)
    assert code == -1
    assert b"gci_radgrd" in lib.gci_last_error()


def test_radgrd_matches_oracle():
    import numpy as np

    import oracle as O
    from grasp_amd import rci

    r, rp, rpor = rci.radgrd(590, 2e-6 / 26, 0.05)
    g = O.Grid(590, 2e-6 / 26, 0.05)
    assert np.array_equal(r, g.r) and np.array_equal(rp, g.rp) and np.array_equal(rpor[1:], g.rpor[1:])


@pytest.mark.parametrize("bad", [float("nan"), float("inf"), -float("inf")])
@pytest.mark.parametrize("where", ["offdiag", "diag"])
def test_upload_rejects_non_finite_values(bad, where):
    # .res rows of a 3x3 block, diagonal last; the check runs before any
    # device call
    import numpy as np

    from grasp_amd import rci

    iendc = np.array([0, 1, 3, 6], dtype=np.int64)
    irow = np.array([1, 1, 2, 1, 2, 3], dtype=np.int32)
    emt = np.array([-1.0, 0.1, -2.0, 0.2, 0.3, -3.0])
    emt[1 if where == "offdiag" else 5] = bad
    with pytest.raises(L.GciError) as e:
        rci.HamiltonianBlock.from_rows(iendc, irow, emt)
    assert e.value.code == L.GCI_ERR_ARG and "non-finite" in str(e.value)
