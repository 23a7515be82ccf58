"""Builds libfinchgpu.so for gfx950 from finch.jl_amd/csrc/*.hip.

Each translation unit is compiled with hipcc in parallel, then linked into one
shared library next to the Python package (finchgpu/libfinchgpu.so), so the
built library travels with the source tree.

Usage: python finch.jl_amd/build.py [--jobs N] [--verbose]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "finchgpu")
LIB = os.path.join(OUT_DIR, "libfinchgpu.so")
# Measurement probe (gather floor), loaded only by the benchmark.
PROBE_SRC = os.path.join(CSRC, "probe")
PROBE_LIB = os.path.join(OUT_DIR, "libfgx_probe.so")
OBJ_DIR = os.path.join(HERE, "_build")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FGX_OFFLOAD_ARCH", "gfx950")

CXXFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # Packed FP32 instructions are kept out of the device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-command-line-argument", f"-I{INCLUDE}",
]


def _compile(src: str, verbose: bool, obj_dir: str, defines) -> str:
    obj = os.path.join(obj_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CXXFLAGS, *[f"-D{d}" for d in defines], "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn and verbose:
        print(warn, file=sys.stderr)
    return obj


def build(jobs: int | None = None, verbose: bool = False, out: str | None = None,
          defines=()) -> str:
    """Compiles every translation unit and links the library.

    `out` and `defines` build a variant elsewhere (tile-shape sweeps, e.g.
    defines=("FGX_SPMV_RUN=16",)); the default builds the in-tree library.
    """
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    lib = LIB if out is None else os.path.join(out, "libfinchgpu.so")
    obj_dir = OBJ_DIR if out is None else os.path.join(out, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, obj_dir, defines), srcs))
    tmp = lib + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if out is None:
        build_probe(verbose)
    return lib


def build_probe(verbose: bool = False) -> str:
    """Builds the benchmark's gather-floor probe library (not the C ABI)."""
    srcs = sorted(glob.glob(os.path.join(PROBE_SRC, "*.hip")))
    tmp = PROBE_LIB + ".tmp"
    cmd = [HIPCC, *CXXFLAGS, "-shared", "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on the probe:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, PROBE_LIB)
    return PROBE_LIB


def clean() -> None:
    shutil.rmtree(OBJ_DIR, ignore_errors=True)
    for f in (LIB, PROBE_LIB):
        if os.path.exists(f):
            os.remove(f)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--clean", action="store_true")
    ap.add_argument("--out", default=None, help="build a variant library into this directory")
    ap.add_argument("-D", dest="defines", action="append", default=[],
                    help="preprocessor definition for a variant build (NAME=VALUE)")
    a = ap.parse_args()
    if a.clean:
        clean()
    print(build(a.jobs, a.verbose, a.out, a.defines))
