"""Build recipe for libdyncore.so (hipcc, gfx950) and the C++ drop-in layer.

The shared libraries are written in-tree under model_amd/_build/ so they travel with the source tree;
nothing is installed into site-packages.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libdyncore.so")
REFAPI_LIB = os.path.join(OUT, "libdyncore_refapi.so")

HIP_SOURCES = ["dyn_capi.hip", "dyn_stencil.hip", "dyn_pressure.hip", "dyn_fused.hip", "dyn_adjoint.hip", "dyn_ensemble.hip",
               "dyn_output.hip", "dyn_turbulence.hip"]
REFAPI_SOURCES = ["refapi/refapi.cpp"]
# the C++ symbols libdyncore_refapi.so defines: what a maintainer localizes in the reference's objects
REFAPI_REPLACED = os.path.join(OUT, "libdyncore_refapi.replaced")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DYN_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # the FP64 plain form keeps the reference's operation order: no FMA contraction
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-variable", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def build(verbose: bool = False) -> str:
    """Compile every HIP source for gfx950 and link libdyncore.so; then the drop-in layer."""
    os.makedirs(OUT, exist_ok=True)
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OUT, os.path.splitext(os.path.basename(src))[0] + ".o")
        objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", "-o", obj, path])
    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    # C++ drop-in layer (reference entry-point names) over the C ABI
    srcs = [os.path.join(CSRC, s) for s in REFAPI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if srcs:
        tmp = REFAPI_LIB + ".tmp"
        _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", "-o", tmp, *srcs,
              f"-L{OUT}", "-ldyncore", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, REFAPI_LIB)
        nm = _run(["nm", "-D", "--defined-only", REFAPI_LIB]).stdout.splitlines()
        syms = sorted({ln.split()[2] for ln in nm if len(ln.split()) == 3 and ln.split()[1] == "T"
                       and ln.split()[2].startswith("_Z")})
        with open(REFAPI_REPLACED + ".tmp", "w") as fh:
            fh.write("\n".join(syms) + "\n")
        os.replace(REFAPI_REPLACED + ".tmp", REFAPI_REPLACED)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
