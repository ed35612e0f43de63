"""Build libsrcwa_amd.so in-tree for gfx950 with hipcc (no JIT, no setup.py).

    python -m stablercwa_amd.build            # compile every source, link the library
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libsrcwa_amd.so")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
# C++ host API driver used by the GPU tests (links the library above)
DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "api_driver.cpp")
DRIVER = os.path.join(OBJ, "api_driver")

SOURCES = ["rcwa1d.hip", "conical.hip", "rcwa2d.hip", "dense.hip", "runtime.cpp", "server.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src: str, obj_dir: str = OBJ, defines=()) -> str:
    obj = os.path.join(obj_dir, src + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + FLAGS + [f"-D{d}" for d in defines] + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build_variant(out_lib: str, defines=(), obj_dir: str = None) -> str:
    """Builds the library with extra preprocessor definitions into out_lib (used by the tuner)."""
    obj_dir = obj_dir or os.path.join(os.path.dirname(out_lib), "_obj")
    os.makedirs(obj_dir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda src: _compile(src, obj_dir, defines), SOURCES))
    r = subprocess.run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", out_lib] + objs + ["-lpthread"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    return out_lib


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if os.path.exists(DRIVER_SRC):
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", INCLUDE, DRIVER_SRC, "-o", DRIVER, "-L", HERE,
               "-lsrcwa_amd", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ API driver failed to build:\n{r.stderr[-4000:]}")
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
