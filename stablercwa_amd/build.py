"""Build libsrcwa_amd.so in-tree for gfx950 with hipcc (no JIT, no setup.py).

    python -m stablercwa_amd.build                 # compile every unit, link the library
    python -m stablercwa_amd.build --incremental   # recompile only units older than their sources

Each family source (rcwa1d.hip, conical.hip, rcwa2d.hip, dense.hip) is compiled as several units: one
"main" unit (its light kernels and host launcher), one for its S-matrix kernel (-DSRCWA_PART_ONLY=1)
(rcwa1d.hip and conical.hip: also their near-field forms, -DSRCWA_PART_ONLY=2) and one per eigen-kernel instantiation QC in
{1, 2, 4, 8, 12, 16} (-DSRCWA_QC_ONLY=QC), so the slow units build in parallel.
"""
from __future__ import annotations

import glob
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
# command-line front end of the JSON request layer (include/stablercwa_amd/rcwa_json.hpp)
JSON_SRC = os.path.join(ROOT, "tests", "cpp", "json_driver.cpp")
JSON_DRIVER = os.path.join(OBJ, "json_driver")

FAMILIES = ["rcwa1d.hip", "conical.hip", "rcwa2d.hip", "dense.hip"]
HOST_SOURCES = ["runtime.cpp", "server.cpp"]
QCS = (1, 2, 4, 8, 12, 16)
# (source, unit): 0 = main, "p1" = the S-matrix part, QC = that eigen instantiation
UNITS = ([(f, 0) for f in FAMILIES] + [(f, "p1") for f in FAMILIES if f != "dense.hip"] + [("rcwa1d.hip", "p2"), ("conical.hip", "p2")] +
         [(f, q) for f in FAMILIES for q in QCS] + [(f, 0) for f in HOST_SOURCES])

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
JOBS = min(16, os.cpu_count() or 8)


def _obj_name(src: str, unit, obj_dir: str) -> str:
    tag = "" if unit == 0 else (f".{unit}" if isinstance(unit, str) else f".q{unit}")
    return os.path.join(obj_dir, src + tag + ".o")


def _unit_defines(unit):
    if unit == 0:
        return []
    if isinstance(unit, str):
        return [f"-DSRCWA_PART_ONLY={unit[1:]}"]
    return [f"-DSRCWA_QC_ONLY={unit}"]


def _newest_input(src: str) -> float:
    # device sources see only csrc/*.hpp; the host sources also include the C-ABI header
    headers = glob.glob(os.path.join(CSRC, "*.hpp"))
    if src in HOST_SOURCES:
        headers += glob.glob(os.path.join(INCLUDE, "*.h"))
    return max(os.path.getmtime(p) for p in [os.path.join(CSRC, src)] + headers)


def _compile(unit, obj_dir: str = OBJ, defines=(), incremental: bool = False) -> str:
    src, part = unit
    obj = _obj_name(src, part, obj_dir)
    if incremental and os.path.exists(obj) and os.path.getmtime(obj) >= _newest_input(src):
        return obj
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    defs = [f"-D{d}" for d in defines] + _unit_defines(part)
    cmd = [HIPCC] + FLAGS + defs + lang + ["-c", os.path.join(CSRC, src), "-o", obj + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src} (unit {part}):\n{r.stderr[-6000:]}")
    os.replace(obj + ".tmp", obj)
    return obj


def _link(objs, out_lib: str) -> None:
    tmp = out_lib + ".tmp"
    r = subprocess.run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, out_lib)


def _compile_all(obj_dir: str, defines=(), incremental: bool = False):
    os.makedirs(obj_dir, exist_ok=True)
    # the slowest units first, so the pool drains evenly
    order = sorted(UNITS, key=lambda u: (u[0] in HOST_SOURCES, isinstance(u[1], int) and u[1] > 0))
    with ThreadPoolExecutor(max_workers=JOBS) as ex:
        done = dict(zip(order, ex.map(lambda u: _compile(u, obj_dir, defines, incremental), order)))
    return [done[u] for u in UNITS]


def build_variant(out_lib: str, defines=(), obj_dir: str = None) -> str:
    """Builds the library with extra preprocessor definitions into out_lib (used by the tuner)."""
    obj_dir = obj_dir or os.path.join(os.path.dirname(out_lib), "_obj")
    _link(_compile_all(obj_dir, defines), out_lib)
    return out_lib


def build(verbose: bool = False, incremental: bool = False) -> str:
    _link(_compile_all(OBJ, incremental=incremental), LIB)
    for src, out in ((DRIVER_SRC, DRIVER), (JSON_SRC, JSON_DRIVER)):
        if not os.path.exists(src):
            continue
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", INCLUDE, src, "-o", out, "-L", HERE,
               "-lsrcwa_amd", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{os.path.basename(src)} failed to build:\n{r.stderr[-4000:]}")
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv[1:])
    sys.exit(0)
