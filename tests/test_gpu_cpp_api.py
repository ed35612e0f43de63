"""The C++17 host API (include/stablercwa_amd/rcwa.hpp) on the GPU, through a small driver program that
builds the C++ Stack / Stack2D types from the fixtures' flat records and calls rcwa::solve, sweep,
solve_conical and solve2d:

* the reference's planar tables at 1e-9 and the compiled reference's vectors at 1e-10
* a sweep equals its single solves bitwise
* the reference's conical Jones tables at 1e-9 max(1, |J|); flux unitarity of a lossless grating
* the reference's crossed-grating tables (DE, s/p amplitudes at 1e-9) and Rayleigh status
* invalid input surfaces as std::invalid_argument with the reference's message; FFT layers as Unsupported
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "stablercwa_amd", "_obj", "api_driver")


def _rec(stack):
    return f"{len(stack)} " + " ".join(repr(float(x)) for x in stack)


def _drive(lines):
    assert os.path.exists(DRIVER), "C++ API driver not built"
    r = subprocess.run([DRIVER], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    blocks, cur = [], []
    for ln in r.stdout.splitlines():
        if ln == "end":
            blocks.append(cur)
            cur = []
        else:
            cur.append(ln)
    assert len(blocks) == len(lines)
    return blocks


def _fields(block):
    head = block[0].split()
    out = {"head": head}
    for ln in block[1:]:
        parts = ln.split()
        out[parts[0]] = np.array([float(x) for x in parts[1:]])
    return out


def test_cpp_planar_tables_and_reference_vectors():
    tables = load_golden("planar_tables.json")["cases"]
    refs = load_golden("planar_ref.json")["cases"]
    cases = tables + refs
    lines = [f"planar {c['pol']} {c['M']} {c['wavelength']!r} {c['theta']!r} {_rec(c['stack'])}" for c in cases]
    for c, b in zip(cases, _drive(lines)):
        f = _fields(b)
        assert f["head"][0] == "ok", (c["name"], b[0])
        tol = 1e-9 if c in tables else 1e-10
        assert np.max(np.abs(f["DEr"] - np.asarray(c["DEr"]))) <= tol, c["name"]
        assert np.max(np.abs(f["DEt"] - np.asarray(c["DEt"]))) <= tol, c["name"]


def test_cpp_sweep_equals_single_solves():
    c = load_golden("planar_ref.json")["cases"][0]
    lams = [1.0, 1.021, 1.053, 1.1]
    pts = " ".join(f"{l!r} {c['theta']!r}" for l in lams)
    lines = [f"sweep {c['pol']} {c['M']} {len(lams)} {pts} {_rec(c['stack'])}"]
    lines += [f"planar {c['pol']} {c['M']} {l!r} {c['theta']!r} {_rec(c['stack'])}" for l in lams]
    blocks = _drive(lines)
    sweep = blocks[0]
    assert sweep[0] == f"sweep {len(lams)}"
    per = [sweep[1 + 5 * i: 6 + 5 * i] for i in range(len(lams))]
    for s, single in zip(per, blocks[1:]):
        assert s == single


def test_cpp_conical_tables():
    cases = load_golden("conical_tables.json")["cases"]
    lines = [f"conical {c['M']} {c['wavelength']!r} {c['theta']!r} {c['phi']!r} {_rec(c['stack'])}" for c in cases]
    blocks = _drive(lines)
    for c, b in zip(cases, blocks):
        f = _fields(b)
        assert f["head"][0] == "ok", (c["id"], b[0])
        for key in ("Jr", "Jt"):
            J = f[key].view(np.complex128)
            R = np.asarray(c[key]).view(np.complex128)
            assert np.max(np.abs(J - R) / np.maximum(1.0, np.abs(R))) <= 1e-9, (c["id"], key)
        # s input: DEr_s + DEr_p of for_angles(90) equals the ss + ps channel sums
        er = f["DEr"].reshape(-1, 4)
        assert np.max(np.abs(f["s_DEr"] - (er[:, 0] + er[:, 2]))) <= 1e-13
    u = _fields(blocks[0])["unitarity"].view(np.complex128)      # binary grating, lossless
    assert np.max(np.abs(u - np.array([1, 0, 0, 1]))) <= 1e-12


def test_cpp_crossed_tables():
    cases = load_golden("rcwa2d_tables.json")["cases"]
    lines = []
    for c in cases:
        j = c["jones"]
        lines.append(f"crossed {c['M1']} {c['M2']} {c['wavelength']!r} {c['theta']!r} {c['phi']!r} "
                     f"{j[0]!r} {j[1]!r} {j[2]!r} {j[3]!r} {c['laurent']} {_rec(c['stack'])}")
    for c, b in zip(cases, _drive(lines)):
        f = _fields(b)
        assert f["head"][0] == "ok", (c["name"], b[0])
        status = int(f["head"][2])
        if c["status"] != 0:
            assert status == 2, c["name"]                 # Status2D::RayleighAnomaly
            continue
        assert status == 0, c["name"]
        tol = lambda ref: c["atol"] + c["rtol"] * np.abs(ref)
        for key in ("DEr", "DEt"):
            ref = np.asarray(c[key])
            assert np.all(np.abs(f[key] - ref) <= tol(ref)), (c["name"], key)
        for key in ("Rs", "Rp", "Ts", "Tp"):
            ref = np.asarray(c[key]).view(np.complex128)
            assert np.all(np.abs(f[key].view(np.complex128) - ref) <= tol(ref)), (c["name"], key)
        assert np.max(np.abs(f["Kx"] - np.asarray(c["Kx"]))) <= 1e-14


def test_cpp_error_conventions():
    good = load_golden("planar_ref.json")["cases"][0]
    bad_edges = list(good["stack"])
    # first layer of the README stack is the grating: period, nI, nII, L, d, kind, K, edges...
    assert bad_edges[7] == 1.0
    bad_edges[11] += 0.01                                  # last edge no longer spans one period
    fft = list(good["stack"])
    K = int(fft[8])
    fft[9 + (K + 1) + 2 * K] = 1.0                         # fourier_method = FFT (automatic samples)
    few = list(fft)
    few[10 + (K + 1) + 2 * K] = 80.0                       # fft_samples = 80 <= 4 M at M = 20
    lines = [f"planar 0 20 {good['wavelength']!r} {good['theta']!r} {_rec(bad_edges)}",
             f"planar 0 20 {good['wavelength']!r} 95.0 {_rec(good['stack'])}",
             f"planar 0 20 {good['wavelength']!r} {good['theta']!r} {_rec(fft)}",
             f"planar 0 20 {good['wavelength']!r} {good['theta']!r} {_rec(few)}"]
    b = _drive(lines)
    assert b[0][0] == "invalid_argument rcwa: grating edges must span exactly one period"
    assert b[1][0] == "invalid_argument rcwa: incidence angle must lie strictly between -90 and 90 degrees"
    assert b[2][0].startswith("ok ")                       # FFT layers are computed
    assert b[3][0] == "invalid_argument rcwa: fft_samples must be at least 4*M+1"


def test_cpp_options_diagnostics(ref_lib):
    """rcwa::solve with Options::diagnostics: Result::max_eig_residual, max_solve_backward_error, n_eig,
    n_dense_solve as the reference's (counts exact), and the timing split filled (rcwa1d.hpp:141-163)."""
    from oracle import ref
    refs = load_golden("planar_ref.json")["cases"][:6]
    lines = [f"planar_diag {c['pol']} {c['M']} {c['wavelength']!r} {c['theta']!r} {_rec(c['stack'])}" for c in refs]
    for c, b in zip(refs, _drive(lines)):
        f = _fields(b)
        assert f["head"][0] == "ok", (c["name"], b[0])
        res, berr, n_eig, n_solve, total, t_eig, t_lin, t_prod = f["diag"]
        o = ref.solve1d_ex(ref_lib, np.asarray(c["stack"]), c["wavelength"], c["theta"], c["M"], c["pol"])
        assert (int(n_eig), int(n_solve)) == (o["n_eig"], o["n_dense_solve"]), c["name"]
        if o["n_eig"]:
            assert 0.0 < res <= 1e-14 and res <= 10 * max(o["max_eig_residual"], 1e-16), (c["name"], res)
        if o["n_dense_solve"]:
            assert 0.0 < berr <= 1e-14, (c["name"], berr)
        assert total > 0.0 and t_eig >= 0.0 and t_lin > 0.0 and t_prod == 0.0
        assert np.max(np.abs(f["DEr"] - np.asarray(c["DEr"]))) <= 1e-10
