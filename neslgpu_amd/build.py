"""Builds neslgpu_amd/libcvl_hip.so, the gfx950 HIP backend of the CVL API,
and neslgpu_amd/bin/vinterp, the VCODE interpreter that drives it.

    python -m neslgpu_amd.build          # compile every source, link

Each csrc/*.hip file is compiled to an object with hipcc (in parallel) and
linked into one shared library that exports the C API of include/cvl.h and
include/cvl_hip_ext.h.  Floating point is compiled without contraction so
that elementwise double arithmetic rounds as the serial CVL's C code does.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcvl_hip.so")
ARCH = os.environ.get("CVLHIP_ARCH", "gfx950")

VCODE = os.path.join(HERE, "vcode")
VINTERP = os.path.join(HERE, "bin", "vinterp")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build_vinterp(verbose: bool = True) -> str:
    """The VCODE interpreter: host C++ only; it loads a CVL library with dlopen."""
    os.makedirs(os.path.dirname(VINTERP), exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(VCODE, "*.cpp")))
    tmp = VINTERP + ".tmp"
    cmd = [CXX, "-std=c++17", "-O2", "-Wall", "-o", tmp, *srcs, "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("building vinterp failed")
    os.replace(tmp, VINTERP)
    if verbose:
        print(f"built {VINTERP}")
    return VINTERP


FUSED_DIR = os.path.join(HERE, "fused")
SUPEROPS_DIR = os.path.join(os.path.dirname(HERE), "tests", "golden", "vcode", "superops")
PROGRAMS_DIR = os.path.join(HERE, "vcode", "programs")


def build_fused(verbose: bool = True) -> list:
    """Fused-operation libraries for the super-op definitions the tests and
    the benchmark use (neslgpu_amd/fused/<program>.so)."""
    from .fusion import codegen
    defs = sorted(glob.glob(os.path.join(SUPEROPS_DIR, "*.superops")) +
                  glob.glob(os.path.join(PROGRAMS_DIR, "*.superops")))
    outs = [os.path.join(FUSED_DIR, os.path.basename(d)[: -len(".superops")] + ".so") for d in defs]
    with ThreadPoolExecutor(max_workers=min(8, max(1, len(defs)))) as ex:
        list(ex.map(lambda a: codegen.build(open(a[0]).read(), a[1], verbose=verbose), zip(defs, outs)))
    return outs


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with ThreadPoolExecutor(max_workers=min(8, len(srcs) + 1)) as ex:
        vint = ex.submit(build_vinterp, verbose)
        objs = list(ex.map(_compile, srcs))
        vint.result()
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    build_fused(verbose)
    return LIB


if __name__ == "__main__":
    build()
