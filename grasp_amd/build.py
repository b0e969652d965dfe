"""Build libgraspci.so (HIP, gfx950) in place with hipcc.

Run as ``python -m grasp_amd.build``.  Every .hip source under csrc/ is
compiled to an object in csrc/_build/ (in parallel), then linked with
rocSOLVER/rocBLAS into grasp_amd/libgraspci.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgraspci.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GCI_OFFLOAD_ARCH", "gfx950")

CXXFLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(HERE, "..", "include"),
]
LDFLAGS = ["-shared", f"--offload-arch={ARCH}", "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas",
           "-Wl,-rpath,/opt/rocm/lib"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str, obj: str):
    cmd = [HIPCC, *CXXFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


SYNTH_DIR = os.path.join(CSRC, "synth")
SYNTH_OUT = os.path.join(HERE, "libgcisynth.so")


def _link(objs, out, extra=()):
    tmp = out + ".tmp"
    cmd = [HIPCC, *LDFLAGS, *objs, *extra, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, out)


def build(verbose: bool = True) -> str:
    """Builds libgraspci.so (the engine) and libgcisynth.so (seeded
    synthetic block generator used by bench.py and tests)."""
    bdir = os.path.join(CSRC, "_build")
    os.makedirs(bdir, exist_ok=True)
    srcs = sources()
    ssrcs = sorted(os.path.join(SYNTH_DIR, f) for f in os.listdir(SYNTH_DIR) if f.endswith(".hip"))
    objs = [os.path.join(bdir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    sobjs = [os.path.join(bdir, "synth_" + os.path.basename(s)[:-4] + ".o") for s in ssrcs]
    jobs = max(1, min(8, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_compile, srcs + ssrcs, objs + sobjs))
    _link(objs, OUT)
    _link(sobjs, SYNTH_OUT, [f"-L{HERE}", "-lgraspci", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print(f"built {OUT} and {SYNTH_OUT}")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
