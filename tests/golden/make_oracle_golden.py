"""Writes tests/golden/oracle_blocks.npz: the oracle's MANEIG eigenvalues for
the seeded blocks that the configuration-size GPU parity tests use.

The oracle (oracle/, the C restatement of MANEIG -> INIEST2 + GDVD(SPICMV2),
rci90/maneig.f90:157-555) takes about 35 s per block on 8 threads, so its
results for these blocks are computed once here and committed; the GPU tests
regenerate each block from its seed (grasp_amd.workloads.clustered_block,
numpy's PCG64 stream, identical on every machine) and check the block's
fingerprint (order, stored entries, sum of values, sum of squared values)
before comparing eigenvalues.

Cases:
  config2_2e4  -- a 2e4-row block at 5% lower fill, 4 eigenpairs (the plain
                  FP64 drop-in at the Fe XVII n <= 6 block size);
  config3_3e4  -- a 3e4-row block at 5% fill, 8 eigenpairs (the precision
                  form's 8-eigenpair, 8-vector solve at a size the oracle
                  finishes).

Run:  python tests/golden/make_oracle_golden.py
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from grasp_amd import workloads as W  # noqa: E402

CASES = {
    "config2_2e4": dict(n=20000, seed=21, fill=0.05, nev=4),
    "config3_3e4": dict(n=30000, seed=31, fill=0.05, nev=8),
}


def block(case):
    c = CASES[case]
    return W.clustered_block(c["n"], c["seed"], fill=c["fill"], sorted_diag=True)


def fingerprint(csr):
    return np.array([csr.n, csr.nnz, float(np.sum(csr.val)), float(np.sum(csr.val * csr.val))])


def main():
    out = {}
    for case, c in CASES.items():
        csr = block(case)
        eav = float(np.mean(csr.val[csr.row_end[1:] - 1]))
        t0 = time.time()
        ev, evec, nl, nm, rc = O.maneig(csr.row_end, csr.col, csr.val, eav, c["nev"],
                                        nthreads=os.cpu_count() or 1)
        assert rc == 0, f"{case}: oracle MANEIG returned {rc}"
        print(f"{case}: n={csr.n} nnz={csr.nnz} eav={eav:.12f} nloops={nl} nmv={nm} "
              f"{time.time() - t0:.1f}s E={ev + eav}")
        out[f"{case}_eval"] = ev
        out[f"{case}_eav"] = np.array([eav])
        out[f"{case}_fingerprint"] = fingerprint(csr)
        out[f"{case}_iters"] = np.array([nl, nm])
    np.savez(os.path.join(HERE, "oracle_blocks.npz"), **out)


if __name__ == "__main__":
    main()
