"""Writes tests/golden/fft_ref.json: the reference's compiled solver (oracle/_ref) on stacks whose grating
layers take FFT-sampled Fourier coefficients (FourierMethod::FFT, fourier.hpp:26-83).

Run in a container that holds the reference tree and has built oracle/_ref (not needed afterwards):

    python tests/golden/make_fft_fixtures.py

Cases: planar TE and TM and conical solves with automatic sample counts (8192, 16384 and 32768 at
M = 6, 70 and 130), power-of-two and other explicit counts (9 at M = 2, 257, 1000, 4096), and gratings whose
edges start at a negative or shifted origin (the phase correction).
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    from oracle import ref
    from stablercwa_amd.types import FourierMethod, Layer, Stack
    from stablercwa_amd.workloads import mirror_stack
    lib = ref.load()
    lib.ref_set_threads(os.cpu_count() or 8)
    FFT = FourierMethod.FFT
    p, f = 0.5747, 0.30
    grating = lambda s: Stack(1.0, [Layer.grating(0.38, [-f * p / 2, f * p / 2, p - f * p / 2], [1.45, 1.0], FFT, s)],
                              1.45, p)
    lossy = Stack(1.2, [Layer.grating(0.25, [-0.3, 0.0125, 0.4, 0.7], [2.3 + 0.4j, 1.1, 3.0 + 0.1j], FFT, 0),
                        Layer.uniform(0.1, 1.7),
                        Layer.grating(0.3, [0.2, 0.45, 1.2], [1.5, 2.2 + 0.05j], FFT, 1000)], 1.5, 1.0)
    mirror = mirror_stack()
    mirror.layers[0] = Layer.grating(0.38, [0.0, 0.3 * 0.574, 0.574], [1.46, 1.0], FFT, 0)
    planar = [("grating_auto", grating(0), 1.053, 66.36, 6), ("grating_257", grating(257), 1.053, 66.36, 6),
              ("grating_9", grating(9), 1.053, 30.0, 2), ("grating_4096", grating(4096), 0.9, 10.0, 12),
              ("lossy_origin", lossy, 0.7, 25.0, 10), ("mirror_M70", mirror, 1.0, 66.36, 70),
              ("mirror_M130", mirror, 1.05, 66.36, 130)]
    out = []
    for name, st, lam, th, M in planar:
        for pol in (0, 1):
            r = ref.solve1d(lib, st.flat(), lam, th, M, pol)
            out.append(dict(kind="planar", name=f"{name}_pol{pol}", stack=st.flat().tolist(), wavelength=lam, theta=th,
                            M=M, pol=pol, DEr=r["DEr"].tolist(), DEt=r["DEt"].tolist(),
                            R=r["R"].view(np.float64).tolist(), T=r["T"].view(np.float64).tolist()))
            print(out[-1]["name"], flush=True)
    conical = [("grating_auto", grating(0), 1.053, 66.36, 17.0, 6), ("lossy_origin", lossy, 0.8, 30.0, 40.0, 8),
               ("grating_1000", grating(1000), 1.0, 20.0, 60.0, 10)]
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    for name, st, lam, th, ph, M in conical:
        N = 2 * M + 1
        f_ = st.flat()
        Jr = np.zeros(8 * N); Jt = np.zeros(8 * N); DEr = np.zeros(4 * N); DEt = np.zeros(4 * N)
        info = ctypes.c_int(); ns = ctypes.c_int()
        assert lib.ref_solve_conical(P(f_), ctypes.c_double(lam), ctypes.c_double(th), ctypes.c_double(ph), M, P(Jr),
                                     P(Jt), P(DEr), P(DEt), ctypes.byref(info), ctypes.byref(ns)) == 0
        out.append(dict(kind="conical", name=name, stack=f_.tolist(), wavelength=lam, theta=th, phi=ph, M=M,
                        DEr=DEr.tolist(), DEt=DEt.tolist(), Jr=Jr.tolist(), eig_info=info.value, n_schur=ns.value))
        print(name, flush=True)
    with open(os.path.join(HERE, "fft_ref.json"), "w") as fh:
        json.dump(dict(source="reference rcwa::solve / solve_conical with FourierMethod::FFT layers (oracle/_ref)",
                       cases=out), fh)
    print(len(out), "cases")


if __name__ == "__main__":
    main()
