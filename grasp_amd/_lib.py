"""ctypes binding of libgraspci.so (the HIP engine's C ABI, include/graspci.h).

The library must exist: there is no CPU fallback.  Build it with
``python -m grasp_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgraspci.so")
_LIB = None

GCI_OK = 0
GCI_ERR_ARG, GCI_ERR_DEVICE, GCI_ERR_NOMEM, GCI_ERR_NOTFOUND, GCI_ERR_NOCONV, GCI_ERR_BUSY = -1, -2, -3, -4, -5, -6
GCI_FP64, GCI_SPLIT32 = 0, 1
GCI_OP_FP64, GCI_OP_SPLIT_BULK, GCI_OP_SPLIT_FULL = 0, 1, 2

d_p = C.POINTER(C.c_double)
i32_p = C.POINTER(C.c_int32)
i64_p = C.POINTER(C.c_int64)


class GciError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"graspci error {code}: {msg}")
        self.code = code


class GciGrid(C.Structure):
    _fields_ = [("n", C.c_int), ("h", C.c_double), ("accy", C.c_double),
                ("r", d_p), ("rp", d_p), ("rpor", d_p)]


class GciOrbitals(C.Structure):
    _fields_ = [("nw", C.c_int), ("ld", C.c_int), ("P", d_p), ("Q", d_p), ("mf", i32_p),
                ("nak", i32_p), ("nkl", i32_p), ("nkj", i32_p)]


class GciPairLists(C.Structure):
    _fields_ = [("npairs", C.c_int64), ("pair_ic", i32_p), ("pair_ir", i32_p), ("one_ia", i32_p),
                ("one_ib", i32_p), ("tsh_off", i64_p), ("tshell", d_p), ("ntshell", C.c_int64),
                ("tb_beg", i64_p), ("tb_end", i64_p), ("label", i32_p), ("coeff", d_p),
                ("nterms", C.c_int64)]


class GciBreitLists(C.Structure):
    _fields_ = [("bt_beg", i64_p), ("bt_end", i64_p), ("label", i32_p), ("coeff", d_p), ("nterms", C.c_int64)]


class GciSpmmShape(C.Structure):
    _fields_ = [("R", C.c_int), ("block", C.c_int), ("nslice", C.c_int), ("depth", C.c_int),
                ("target", C.c_int64)]


class GciDvdParams(C.Structure):
    _fields_ = [("nvec", C.c_int), ("lim", C.c_int), ("mblock", C.c_int), ("maxiter", C.c_int),
                ("crite", C.c_double), ("critc", C.c_double), ("critr", C.c_double),
                ("niv", C.c_int), ("mode", C.c_int), ("refine", C.c_int)]


class GciDvdStats(C.Structure):
    _fields_ = [("nloops", C.c_int), ("nmv", C.c_int), ("ierr", C.c_int), ("refine_loops", C.c_int),
                ("seconds", C.c_double)]


# every symbol include/graspci.h declares
EXPORTS = [
    "gci_last_error", "gci_init", "gci_radgrd", "gci_rk_table", "gci_one_body", "gci_h_assemble",
    "gci_hmat_upload", "gci_hmat_free", "gci_hmat_info", "gci_hmat_download", "gci_spmm",
    "gci_spmm_device", "gci_davidson", "gci_maneig", "gci_batch_solve", "gci_server_create",
    "gci_server_destroy", "gci_server_submit", "gci_server_poll", "gci_server_drain",
    "gci_profile_enable", "gci_profile_read", "gci_spmm_get_shape", "gci_spmm_set_shape", "gci_spmm_time",
    "gci_spmm_autotune",
    "gci_spmm_layout",
    "gci_server_submit_shared",
    "gci_host_alloc", "gci_host_free",
    "gci_spmm_device_noalloc",
    "gci_yzk", "gci_maneig_mcdhf",
    "gci_breit_create", "gci_breit_table", "gci_breit_integral", "gci_breit_free", "gci_h_assemble_tb",
    "gci_hstream_begin", "gci_hstream_push", "gci_hstream_end", "gci_hstream_free",
]
GCI_NOALLOC_SCRATCH_BYTES = 48 << 20


def lib():
    """Load libgraspci.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run python -m grasp_amd.build")
        L = C.CDLL(LIB_PATH)
        L.gci_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def check(code: int):
    if code != GCI_OK:
        raise GciError(code, lib().gci_last_error().decode(errors="replace"))
    return code


def ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def dp(a):
    return ptr(a, d_p)


def i32p(a):
    return ptr(a, i32_p)


def i64p(a):
    return ptr(a, i64_p)
