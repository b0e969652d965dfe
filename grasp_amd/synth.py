"""Seeded synthetic CI blocks generated on the device (libgcisynth.so).

Input generation for benchmarks and large tests: a block is a pure function
of its parameters (counter-based hashing), built directly in HBM, so the
workloads of the batched and precision forms (10^4..10^5-row blocks,
10^8..10^9 stored elements) do not pass through host memory.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import _lib as L
from .rci import HamiltonianBlock

_SYN = None


def lib():
    global _SYN
    if _SYN is None:
        L.lib()  # libgraspci first (libgcisynth links against it)
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgcisynth.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} not built: run python -m grasp_amd.build")
        _SYN = C.CDLL(path)
    return _SYN


def synth_block(n: int, seed: int, fill: float = 0.05, diag_mean: float = -338.0, diag_sd: float = 5.0,
                off_sd: float = 1.0e-2, group=(4, 48), inner: float = 0.6, precision: int = L.GCI_SPLIT32,
                shift_eav: bool = True, sorted_diag: bool = True):
    """Returns (HamiltonianBlock, EAV).  With shift_eav the diagonal is held
    as H_ii - EAV, as MANEIG loads it (rci90/maneig.f90:326); sorted_diag
    orders the diagonal ascending, as a CSF list that starts from the
    reference configuration is ordered (INIEST2's leading block then holds
    the lowest states)."""
    h = C.c_void_p()
    eav = C.c_double(0.0)
    L.check(lib().gcs_synth_block(C.c_int64(int(n)), C.c_uint64(int(seed)), C.c_double(fill),
                                  C.c_double(diag_mean), C.c_double(diag_sd), C.c_double(off_sd),
                                  C.c_int(group[0]), C.c_int(group[1]), C.c_double(inner), C.c_int(precision),
                                  C.c_int(int(shift_eav)), C.c_int(int(sorted_diag)), C.byref(eav),
                                  C.byref(h)))
    nn, nz, pr = C.c_int64(), C.c_int64(), C.c_int()
    L.check(L.lib().gci_hmat_info(h, C.byref(nn), C.byref(nz), C.byref(pr)))
    return HamiltonianBlock(h, nn.value, nz.value, pr.value), eav.value


def batched_workload(nblocks: int = 64, seed: int = 2024, lo: float = 1.0e2, hi: float = 5.0e4,
                     fill_lo: float = 0.01, fill_hi: float = 0.10):
    """The batched configuration: block sizes log-uniform in [lo, hi], fill
    uniform in [fill_lo, fill_hi], all seeded.  Returns [(n, fill, seed_b)]."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(math.log(lo), math.log(hi), size=nblocks))).astype(np.int64)
    fills = rng.uniform(fill_lo, fill_hi, size=nblocks)
    seeds = rng.integers(1, 2**62, size=nblocks)
    return [(int(n), float(f), int(s)) for n, f, s in zip(sizes, fills, seeds)]
