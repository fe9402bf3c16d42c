"""Build libssf_gpu.so (gfx950) in-tree with hipcc.

Each source is compiled to its own object in ssf_amd/_build (in parallel), then
the objects are linked. incremental=True recompiles only the sources whose
object is older than the source or any header of csrc/ and include/.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libssf_gpu.so")
SOURCES = ["ssf_kernels.hip", "ssf_solver.hip", "ssf_elast.hip", "ssf_api.hip", "ssf_batch.hip",
           "ssf_batch_api.hip", "ssf_serve.hip", "ssf_server.hip", "ssf_layout.hip",
           "ssf_post.hip", "ssf_pattern.hip", "ssf_bicg.hip", "ssf_ildl.hip", "ssf_nr.hip",
           "ssf_renumber.hip", "ssf_codenum.hip", "ssf_noalloc.hip", "ssf_assemble.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are disabled: hipcc forms them itself from float code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _headers_mtime() -> float:
    hs = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    return max(os.path.getmtime(h) for h in hs)


def build(verbose: bool = False, incremental: bool = False, jobs: int = 8) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    hdr = _headers_mtime()

    def compile_one(src: str) -> str:
        s = os.path.join(CSRC, src)
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        if incremental and os.path.exists(o) and os.path.getmtime(o) >= max(os.path.getmtime(s), hdr):
            return o
        cmd = [hipcc, *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c", s, "-o", o + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
        if out.strip():
            print(out, flush=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}")
        os.replace(o + ".tmp", o)
        return o

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB




DRIVER = os.path.join(ROOT, "tests", "cxx", "gmatrix_driver")


def build_driver(verbose: bool = False) -> str:
    """The host C++ caller of the boundary (tests/cxx/gmatrix_driver.cpp with the
    gmatrix_gpu mirror of ssf_amd/host), linked against libssf_gpu.so."""
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(HERE, "host"), os.path.join(HERE, "host", "gmatrix_gpu.cpp"),
           os.path.join(ROOT, "tests", "cxx", "gmatrix_driver.cpp"), "-o", DRIVER + ".tmp",
           "-L", HERE, "-lssf_gpu", "-Wl,-rpath,$ORIGIN/../../ssf_amd",
           "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose="-v" in sys.argv, incremental="-i" in sys.argv)
    build_driver(verbose="-v" in sys.argv)
    print("built", LIB)
    sys.exit(0)
