"""Build recipe for the mdz libraries.

* ``compression_amd/libmdz.so``: every ``csrc/*.hip`` compiled for gfx950 with
  hipcc, in parallel, into ``compression_amd/_build/`` and linked in-tree;
* ``compression_amd/libmdz_host.so``: the host encoders and workload
  generators (``host/*.c``, gcc, zlib).
Both are built in-tree so they travel with the source tree.  Usage:
``python -m compression_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmdz.so")
HOST_LIB = os.path.join(HERE, "libmdz_host.so")
CC = os.environ.get("CC", "gcc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CXXFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed-FP32 VALU ops are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *CXXFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_host(verbose: bool = False) -> str:
    srcs = sorted(glob.glob(os.path.join(HERE, "host", "*.c")))
    tmp = HOST_LIB + ".tmp"
    cmd = [CC, "-O2", "-g", "-fPIC", "-std=gnu11", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-pthread", f"-I{os.path.join(ROOT, 'include')}", "-shared", "-o", tmp, *srcs, "-lz", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host library build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, HOST_LIB)
    if verbose:
        print(f"built {HOST_LIB} from {len(srcs)} sources")
    return HOST_LIB


def build(verbose: bool = False) -> str:
    build_host(verbose)
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
