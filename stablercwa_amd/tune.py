"""Autotuner for the eigen and S-matrix kernels' compile-time constants (stablercwa_amd/csrc/tuned.hpp).

Two search spaces (--space):

* eig: the eigensolver dominates every workload; its speed depends on how far each column prefetches in
  the QR sweep (SRCWA_EIG_RING), on the occupancy the workgroup kernels are built for (SRCWA_EIG_OCC:
  fewer registers and more resident matrices per CU, against spills) and on the order up to which one
  wave solves a matrix alone (SRCWA_WAVE_EIG_MAX_N); scored by the summed eigen-stage time;
* smat: the occupancy the S-matrix kernels are built for (SRCWA_SMAT_OCC: one instance per SIMD with no
  spills, or two with a few spills around phase boundaries); scored per family by its S-matrix stage.

The best point differs between chips and compilers, so it is measured, not guessed:

    python -m stablercwa_amd.tune build [--space S]              # CPU: one library per variant in _tune/
    python -m stablercwa_amd.tune measure [--space S] -o r.json  # GPU: time the stages of each variant
    python -m stablercwa_amd.tune write [--space S] r.json       # rewrite csrc/tuned.hpp with the fastest

The measured workloads are the benchmark's: the planar M = 60 sweep (8192 solves), 512 instances of
the conical batch and 128 of the crossed pillar batch.  Every variant's planar results are also checked
against the first variant's (max |dDE| <= 1e-10).
"""
from __future__ import annotations

import argparse
import ctypes
import itertools
import json
import os
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TUNE_DIR = os.path.join(HERE, "_tune")
TUNED_HPP = os.path.join(HERE, "csrc", "tuned.hpp")

RINGS = (2, 4, 8)
OCCS = (0, 3, 4, 5)
WAVE_MAX = (64,)


SMAT_OCCS = (2, 1)


def variants(space="eig"):
    if space == "smat":
        for occ in SMAT_OCCS:
            yield f"s{occ}", dict(SRCWA_SMAT_OCC=occ)
        return
    for ring, occ, wmax in itertools.product(RINGS, OCCS, WAVE_MAX):
        yield f"r{ring}_o{occ}_w{wmax}", dict(SRCWA_EIG_RING=ring, SRCWA_EIG_OCC=occ, SRCWA_WAVE_EIG_MAX_N=wmax)


def cmd_build(args):
    from . import build as b
    for name, defs in variants(args.space):
        out = os.path.join(TUNE_DIR, name, "libsrcwa_amd.so")
        if os.path.exists(out) and not args.force:
            continue
        os.makedirs(os.path.dirname(out), exist_ok=True)
        t0 = time.time()
        b.build_variant(out, [f"{k}={v}" for k, v in defs.items()])
        print(f"built {name} in {time.time() - t0:.0f} s", flush=True)


def _load(path):
    from . import _lib
    lib = ctypes.CDLL(path)
    for name, (res, argt) in _lib.PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, argt
    return lib


def _plans(lib, ctx):
    """The three workloads as raw C-ABI plans on the variant library."""
    from . import _lib, workloads
    from .rcwa import _pack_stacks
    ptr = _lib.ptr
    out = {}
    stacks, probs = workloads.mirror_sweep()
    flat, offs = _pack_stacks(stacks)
    P = len(probs)
    idx = np.zeros(P, np.int32)
    wl = np.array([p[1] for p in probs]); th = np.array([p[2] for p in probs])
    M = np.array([p[3] for p in probs], np.int32); pol = np.array([p[4] for p in probs], np.int32)
    h = ctypes.c_void_p()
    _lib.check(lib.srcwa_plan1d_create(ctx, 1, ptr(flat), ptr(offs), P, ptr(idx), ptr(wl), ptr(th), ptr(M), ptr(pol),
                                       0, ctypes.byref(h)))
    out["planar"] = (h, P, (flat, offs, idx, wl, th, M, pol))
    cp = workloads.conical_batch(512)
    flat, offs = _pack_stacks([p.stack for p in cp])
    idx = np.arange(len(cp), dtype=np.int32)
    wl = np.array([p.wavelength for p in cp]); th = np.array([p.theta for p in cp]); ph = np.array([p.phi for p in cp])
    M = np.array([p.M for p in cp], np.int32)
    h = ctypes.c_void_p()
    _lib.check(lib.srcwa_plan_conical_create(ctx, len(cp), ptr(flat), ptr(offs), len(cp), ptr(idx), ptr(wl), ptr(th),
                                             ptr(ph), ptr(M), 0, ctypes.byref(h)))
    out["conical"] = (h, len(cp), (flat, offs, idx, wl, th, ph, M))
    xp = workloads.pillar_batch(128)
    flat, offs = _pack_stacks([p.stack for p in xp])
    idx = np.arange(len(xp), dtype=np.int32)
    wl = np.array([p.wavelength for p in xp]); th = np.array([p.theta for p in xp]); ph = np.array([p.phi for p in xp])
    M1 = np.array([p.M1 for p in xp], np.int32); M2 = np.array([p.M2 for p in xp], np.int32)
    jones = np.tile([0.6, 0.0, 0.8, 0.0], len(xp))
    h = ctypes.c_void_p()
    _lib.check(lib.srcwa_plan2d_create(ctx, len(xp), ptr(flat), ptr(offs), len(xp), ptr(idx), ptr(wl), ptr(th),
                                       ptr(ph), ptr(M1), ptr(M2), ptr(jones), None, 0, ctypes.byref(h)))
    out["crossed"] = (h, len(xp), (flat, offs, idx, wl, th, ph, M1, M2, jones))
    return out


def _measure_one(path):
    from . import _lib
    lib = _load(path)
    ctx = ctypes.c_void_p()
    _lib.check(lib.srcwa_context_create(0, ctypes.byref(ctx)))
    plans = _plans(lib, ctx)
    res = {}
    planar_der = None
    for name, (h, n, keep) in plans.items():
        _lib.check(lib.srcwa_plan_execute(h))
        _lib.check(lib.srcwa_plan_sync(h))
        _lib.check(lib.srcwa_plan_execute(h))
        _lib.check(lib.srcwa_plan_sync(h))
        ms = np.zeros(4)
        _lib.check(lib.srcwa_plan_stage_ms(h, _lib.ptr(ms)))
        res[name] = dict(eig_ms=float(ms[1]), smat_ms=float(ms[2]), total_ms=float(ms[3]), stages_ms=ms.tolist())
        if name == "planar":
            o = lib.srcwa_plan_orders(h)
            planar_der = np.empty(o)
            _lib.check(lib.srcwa_plan_results(h, None, None, _lib.ptr(planar_der), None, None, None))
        lib.srcwa_plan_destroy(h)
    lib.srcwa_context_destroy(ctx)
    return res, planar_der


def cmd_measure(args):
    rows = {}
    ref = None
    names = [n for n, _ in variants(args.space)]
    for name in names:
        path = os.path.join(TUNE_DIR, name, "libsrcwa_amd.so")
        if not os.path.exists(path):
            continue
        # each variant in its own process: two libraries of the same symbols must not share one
        import subprocess
        r = subprocess.run([sys.executable, "-m", "stablercwa_amd.tune", "_one", path, args.output + "." + name],
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            print(f"{name}: failed\n{r.stderr[-2000:]}", flush=True)
            continue
        with open(args.output + "." + name) as f:
            row = json.load(f)
        der = np.array(row.pop("planar_der"))
        if ref is None:
            ref = der
        row["max_dDE_vs_first"] = float(np.max(np.abs(der - ref)))
        stage = "smat_ms" if args.space == "smat" else "eig_ms"
        row["score_ms"] = sum(row[k][stage] for k in ("planar", "conical", "crossed"))
        rows[name] = row
        print(name, json.dumps(row), flush=True)
    with open(args.output, "w") as f:
        json.dump(rows, f, indent=1)


def cmd_one(args):
    res, der = _measure_one(args.path)
    res["planar_der"] = der.tolist()
    with open(args.out, "w") as f:
        json.dump(res, f)


def cmd_write(args):
    """Ring depth and wave-size limit by the total eigen-stage time; then, at that ring depth, each
    family's occupancy by its own eigen-stage time (the families' kernels are built independently)."""
    with open(args.results) as f:
        rows = json.load(f)
    ok = {k: v for k, v in rows.items() if v["max_dDE_vs_first"] <= 1e-10}
    if args.space == "smat":
        defs = dict(variants("smat"))
        chosen = {}
        for fam, key in (("planar", "SRCWA_SMAT_OCC_PLANAR"), ("conical", "SRCWA_SMAT_OCC_CONICAL"),
                         ("crossed", "SRCWA_SMAT_OCC_CROSSED")):
            chosen[key] = defs[min(ok, key=lambda k: ok[k][fam]["smat_ms"])]["SRCWA_SMAT_OCC"]
        _write_tuned(chosen)
        return
    best = min(ok, key=lambda k: ok[k]["score_ms"])
    defs = dict(variants())
    chosen = {"SRCWA_EIG_RING": defs[best]["SRCWA_EIG_RING"], "SRCWA_WAVE_EIG_MAX_N": defs[best]["SRCWA_WAVE_EIG_MAX_N"]}
    same = [k for k in ok if all(defs[k][d] == v for d, v in chosen.items())]
    for fam, key in (("planar", "SRCWA_EIG_OCC_PLANAR"), ("conical", "SRCWA_EIG_OCC_CONICAL"),
                     ("crossed", "SRCWA_EIG_OCC_CROSSED")):
        chosen[key] = defs[min(same, key=lambda k: ok[k][fam]["eig_ms"])]["SRCWA_EIG_OCC"]
    _write_tuned(chosen)


def _write_tuned(chosen):
    src = open(TUNED_HPP).read()
    for k, v in chosen.items():
        src = re.sub(rf"(#define {k} )\d+", rf"\g<1>{v}", src)
    with open(TUNED_HPP, "w") as f:
        f.write(src)
    print(f"chosen {chosen}; wrote {TUNED_HPP}")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build"); b.add_argument("--force", action="store_true")
    m = sub.add_parser("measure"); m.add_argument("-o", "--output", default="tune_results.json")
    o = sub.add_parser("_one"); o.add_argument("path"); o.add_argument("out")
    w = sub.add_parser("write"); w.add_argument("results")
    for p in (b, m, w):
        p.add_argument("--space", choices=("eig", "smat"), default="eig")
    args = ap.parse_args()
    {"build": cmd_build, "measure": cmd_measure, "_one": cmd_one, "write": cmd_write}[args.cmd](args)


if __name__ == "__main__":
    main()
