"""Build libaurora_hip.so (gfx950) in-tree.

    python aurora-pic_amd/build.py

Compiles every csrc/*.hip with hipcc for gfx950 only and links the C-ABI
shared library into aurora-pic_amd/lib/.  Objects go to aurora-pic_amd/_build/.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
SO = os.path.join(LIB, "libaurora_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++20", "--offload-arch=gfx950", "-fPIC",
    # packed FP32 instructions are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # contract a*b+c only within one source expression: the same expression
    # rounds the same way in every kernel that inlines it
    "-ffp-contract=on",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{os.path.join(REPO, 'include')}", f"-I{CSRC}",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


HOST = os.path.join(HERE, "host")
HOST_SO = os.path.join(LIB, "libaurora_gpu.so")
HOST_TESTS = os.path.join(LIB, "aurora_host_tests")
CXX = os.environ.get("CXX", "g++")


def build_host() -> str:
    """C++20 host classes (host/aurora_gpu.cpp, host/particle_state.cpp, host/diagnostics.cpp) over the C ABI, and their
    test program; plain g++, linked against libaurora_hip.so."""
    inc = [f"-I{os.path.join(REPO, 'include')}", f"-I{HOST}"]
    rpath = "-Wl,-rpath,$ORIGIN"
    cmd = [CXX, "-std=c++20", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", *inc,
           os.path.join(HOST, "aurora_gpu.cpp"), os.path.join(HOST, "particle_state.cpp"),
           os.path.join(HOST, "diagnostics.cpp"),
           "-o", HOST_SO + ".tmp", f"-L{LIB}",
           "-laurora_hip", rpath]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host library build failed:\n{r.stderr[-6000:]}")
    os.replace(HOST_SO + ".tmp", HOST_SO)
    cmd = [CXX, "-std=c++20", "-O2", "-Wall", "-Wextra", *inc,
           os.path.join(HOST, "tests", "host_tests.cpp"), "-o", HOST_TESTS + ".tmp", f"-L{LIB}",
           "-laurora_gpu", "-laurora_hip", rpath]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host tests build failed:\n{r.stderr[-6000:]}")
    os.replace(HOST_TESTS + ".tmp", HOST_TESTS)
    return HOST_SO


def build(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(_compile, sources))
    tmp = SO + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, SO)
    build_host()
    return SO


if __name__ == "__main__":
    print(build(int(sys.argv[1]) if len(sys.argv) > 1 else 8))
