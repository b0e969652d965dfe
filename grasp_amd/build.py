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


FORT_DIR = os.path.join(HERE, "fortran")
FORT_BUILD = os.path.join(FORT_DIR, "_build")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")


def build_fortran(verbose: bool = True) -> bool:
    """Builds the Fortran host layer (ISO_C_BINDING module gci_bind, the
    rci90-shaped wrappers gci_rci) into fortran/_build/libgcifort.so and the
    example host program fortran/_build/gci_drive.  Returns False (and
    builds nothing) when no Fortran compiler is installed."""
    if not os.path.exists(FC):
        if verbose:
            print(f"build_fortran: {FC} not found; Fortran host layer not built")
        return False
    os.makedirs(FORT_BUILD, exist_ok=True)
    fflags = ["-O2", "-fPIC", "-module-dir", FORT_BUILD]
    link = [f"-L{HERE}", "-lgraspci", "-Wl,-rpath,$ORIGIN/../.."]
    objs = []
    for name in ("gci_bind", "gci_rci"):
        obj = os.path.join(FORT_BUILD, name + ".o")
        cmd = [FC, *fflags, "-c", os.path.join(FORT_DIR, name + ".f90"), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{FC} failed for {name}.f90:\n{r.stderr[-4000:]}")
        objs.append(obj)
    steps = [[FC, "-shared", *objs, *link, "-o", os.path.join(FORT_BUILD, "libgcifort.so")],
             [FC, *fflags, os.path.join(FORT_DIR, "gci_drive.f90"), *objs, *link, "-o",
              os.path.join(FORT_BUILD, "gci_drive")]]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran link failed:\n{r.stderr[-4000:]}")
    if verbose:
        print(f"built {FORT_BUILD}/libgcifort.so and gci_drive")
    return True


def build(verbose: bool = True) -> str:
    """Builds libgraspci.so (the engine), libgcisynth.so (seeded synthetic
    block generator used by bench.py and tests) and the Fortran host layer."""
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
    build_fortran(verbose)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
