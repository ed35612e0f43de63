"""Writes tests/golden/rcwa2d_tables.json and tests/golden/rcwa2d_ref.json.

Run in a container that holds the reference tree (not needed afterwards):

    python tests/golden/make_rcwa2d_fixtures.py [/root/reference]

rcwa2d_tables.json -- the reference's 24 frozen crossed-grating cases (tests/rcwa2d_reference.inc:
    far-field arrays, expected Rayleigh failures, layer operators), dumped with their stacks by
    dump_rcwa2d_ref.cpp compiled against the reference headers.
rcwa2d_ref.json    -- the reference's compiled solver (oracle/_ref) on seeded circular-pillar instances
    drawn like the benchmark's crossed-grating batch at small truncation orders, plus normal-incidence
    C4-symmetric cases (degenerate eigenvalues).
"""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main(ref_root="/root/reference"):
    exe = "/tmp/srcwa_dump_rcwa2d_ref"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ref_root}/include", f"-I{ref_root}/tests",
                    os.path.join(HERE, "dump_rcwa2d_ref.cpp"), "-o", exe], check=True)
    tables = json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)
    with open(os.path.join(HERE, "rcwa2d_tables.json"), "w") as f:
        json.dump(tables, f)

    from oracle import ref
    from stablercwa_amd.workloads import pillar_batch
    lib = ref.load()
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    out = []
    probs = pillar_batch(10, seed=0xC41, orders=(2, 3))
    probs += [p for p in pillar_batch(4, seed=0xC42, orders=(3,), oblique_fraction=0.0)]
    for i, p in enumerate(probs):
        f = p.stack.flat()
        N = (2 * p.M1 + 1) * (2 * p.M2 + 1)
        jones = np.array([np.sin(np.radians(50.0)), 0.0, np.cos(np.radians(50.0)), 0.0])
        DEr = np.zeros(N); DEt = np.zeros(N)
        Rs = np.zeros(2 * N); Rp = np.zeros(2 * N); Ts = np.zeros(2 * N); Tp = np.zeros(2 * N)
        st = ctypes.c_int(); info = ctypes.c_int()
        rc = lib.ref_solve2d(P(f), ctypes.c_double(p.wavelength), ctypes.c_double(p.theta), ctypes.c_double(p.phi),
                             p.M1, p.M2, P(jones), P(DEr), P(DEt), P(Rs), P(Rp), P(Ts), P(Tp), ctypes.byref(st),
                             ctypes.byref(info))
        assert rc == 0, lib.ref_last_error()
        out.append(dict(name=f"pillar_{i}", stack=f.tolist(), wavelength=p.wavelength, theta=p.theta, phi=p.phi,
                        M1=p.M1, M2=p.M2, jones=jones.tolist(), status=st.value, eig_info=info.value,
                        DEr=DEr.tolist(), DEt=DEt.tolist(), Rs=Rs.tolist(), Rp=Rp.tolist(), Ts=Ts.tolist(),
                        Tp=Tp.tolist()))
    with open(os.path.join(HERE, "rcwa2d_ref.json"), "w") as f:
        json.dump(dict(source="reference rcwa::solve2d built from its headers (oracle/_ref)", cases=out), f)
    print(len(tables["cases"]), "tables;", len(out), "reference vectors; statuses", [c["status"] for c in out])


if __name__ == "__main__":
    main(*sys.argv[1:])
