"""Builds openntc_amd/libopenntc_amd.so for gfx950.

HIP sources (*.hip) go through hipcc with --offload-arch=gfx950; host sources
(*.cpp) through g++ against the HIP runtime headers. Both with
-ffp-contract=off: every fused multiply-add in the library is an explicit fma.
Usage: python -m openntc_amd.build [-j N] [--incremental]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libopenntc_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SOURCES = ["kernels_plain.hip", "kernels_grid.hip", "kernels_tex.hip", "kernels_fused.hip", "kernels_fused_pair.hip", "kernels_split.hip", "kernels_eval.hip", "kernels_eval_mfma.hip"]
CXX_SOURCES = ["context.cpp", "runtime.cpp", "capi.cpp", "synthetic.cpp", "batch.cpp", "manifest.cpp", "images.cpp"]

INCLUDES = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-munsafe-fp-atomics",
    # no packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
CXX_FLAGS = [
    "-O2", "-fPIC", "-std=c++20", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
    "-I", os.path.join(ROCM, "include"), "-Wall", "-Wno-unused-result", "-Wno-unused-function",
]


def _newest_header() -> float:
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    for d, _, fs in os.walk(os.path.join(ROOT, "include")):
        hdrs += [os.path.join(d, f) for f in fs if f.endswith(".h")]
    return max(os.path.getmtime(h) for h in hdrs)


def _compile(src: str, incremental: bool = False) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if incremental and os.path.exists(obj):
        t = os.path.getmtime(obj)
        if t > os.path.getmtime(path) and t > _newest_header():
            return obj
    if src.endswith(".hip"):
        cmd = [os.path.join(ROCM, "bin", "hipcc"), *HIPCC_FLAGS, *INCLUDES, "-c", path, "-o", obj]
    else:
        cmd = ["g++", *CXX_FLAGS, *INCLUDES, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 8, verbose: bool = True, incremental: bool = False) -> str:
    """Compiles every source (incremental: only those older than their
    object or than any header) and links the library."""
    os.makedirs(OBJ, exist_ok=True)
    sources = [s for s in HIP_SOURCES + CXX_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), sources))
    tmp = LIB + ".tmp"
    cmd = [os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp,
           *objs, "-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_api_flow(verbose)
    return LIB


API_FLOW_SRC = os.path.join(ROOT, "tests", "cpp", "api_flow.cpp")


def build_api_flow(verbose: bool = True) -> str | None:
    """The C++ caller of the public API (tests/cpp/api_flow.cpp), linked
    against the library the way an application links it; found at run time
    relative to its own directory."""
    if not os.path.exists(API_FLOW_SRC):
        return None
    out = API_FLOW_SRC[:-len(".cpp")]
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-D__HIP_PLATFORM_AMD__", *INCLUDES, "-I",
           os.path.join(ROCM, "include"), API_FLOW_SRC, "-o", out + ".tmp", "-L", HERE, "-lopenntc_amd",
           "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../openntc_amd",
           "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"api_flow build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--incremental", action="store_true")
    a = ap.parse_args()
    try:
        build(a.j, incremental=a.incremental)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
