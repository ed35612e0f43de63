"""Pins the crossed-grating oracle (oracle/rcwa2d.py).

* the reference's 24 frozen cases (tests/rcwa2d_reference.inc): far-field arrays at atol + rtol |b| = 1e-9,
  layer operators Eps, Cx, Cy at 1e-12, expected Rayleigh failures
* the reference's compiled solver on seeded circular-pillar instances, 1e-10
* a y-invariant crossed grating equals the planar solve (TE, TM)
* lossless energy conservation
* a planted mistake (Cx and Cy swapped) fails the tables
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle import flat, rcwa1d, rcwa2d


def _c(a):
    return np.asarray(a).view(np.complex128)


CASES = load_golden("rcwa2d_tables.json")["cases"]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['id']}_{c['name']}")
def test_oracle_rcwa2d_tables(case):
    j = case["jones"]
    r = rcwa2d.solve(np.array(case["stack"]), case["wavelength"], case["theta"], case["phi"], case["M1"], case["M2"],
                     (complex(j[0], j[1]), complex(j[2], j[3])), laurent=bool(case["laurent"]))
    if case["status"] != 0:
        assert r["status"] == 2          # Rayleigh anomaly
        return
    assert r["status"] == 0
    tol = lambda b: case["atol"] + case["rtol"] * np.abs(b)
    for key in ("DEr", "DEt"):
        assert np.all(np.abs(r[key] - np.asarray(case[key])) <= tol(np.asarray(case[key]))), key
    for key in ("Rs", "Rp", "Ts", "Tp", "Ex_r", "Ey_r", "Ex_t", "Ey_t"):
        ref = _c(case[key])
        assert np.all(np.abs(r[key] - ref) <= tol(ref)), key


@pytest.mark.parametrize("case", [c for c in CASES if c["operators"]], ids=lambda c: f"{c['id']}_{c['name']}")
def test_oracle_layer_operators(case):
    st = flat.decode2d(np.array(case["stack"]))
    for op in case["operators"]:
        layer = st.layers[op["layer"]]
        eps, cx, cy = rcwa2d.layer_operators(layer, st.Lx, st.Ly, case["M1"], case["M2"])
        N = op["N"]
        for mine, key in ((eps, "eps"), (cx, "cx"), (cy, "cy")):
            ref = _c(op[key]).reshape(N, N)           # stored row by row
            assert np.max(np.abs(mine - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref))), key


@pytest.mark.parametrize("case", load_golden("rcwa2d_ref.json")["cases"], ids=lambda c: c["name"])
def test_oracle_rcwa2d_reference_build(case):
    j = case["jones"]
    r = rcwa2d.solve(np.array(case["stack"]), case["wavelength"], case["theta"], case["phi"], case["M1"], case["M2"],
                     (complex(j[0], j[1]), complex(j[2], j[3])))
    assert r["status"] == 0
    assert np.max(np.abs(r["DEr"] - np.asarray(case["DEr"]))) <= 1e-10
    assert np.max(np.abs(r["DEt"] - np.asarray(case["DEt"]))) <= 1e-10


def test_y_invariant_equals_planar():
    from stablercwa_amd.types import Layer, Layer2D, Stack, Stack2D, XProfile2D
    edges = [0.0, 0.21, 0.47, 0.9]
    vals = [1.7, 1.0, 2.1]
    s2 = Stack2D(1.0, 1.5, [Layer2D.patterned(0.3, [0.0, 0.7], [XProfile2D(edges, vals)])], 0.9, 0.7)
    s1 = Stack(1.0, [Layer.grating(0.3, edges, vals)], 1.5, 0.9)
    M = 8
    te = rcwa1d.solve(s1.flat(), 0.83, 27.0, M, 0)
    tm = rcwa1d.solve(s1.flat(), 0.83, 27.0, M, 1)
    r_te = rcwa2d.solve(s2.flat(), 0.83, 27.0, 0.0, M, 0, (1.0, 0.0))
    r_tm = rcwa2d.solve(s2.flat(), 0.83, 27.0, 0.0, M, 0, (0.0, 1.0))
    assert np.max(np.abs(r_te["DEr"] - te["DEr"])) <= 1e-11 and np.max(np.abs(r_te["DEt"] - te["DEt"])) <= 1e-11
    assert np.max(np.abs(r_tm["DEr"] - tm["DEr"])) <= 1e-11 and np.max(np.abs(r_tm["DEt"] - tm["DEt"])) <= 1e-11


def test_lossless_energy():
    case = [c for c in CASES if c["name"] == "pillar_sq"][0]
    j = case["jones"]
    r = rcwa2d.solve(np.array(case["stack"]), case["wavelength"], case["theta"], case["phi"], case["M1"], case["M2"],
                     (complex(j[0], j[1]), complex(j[2], j[3])))
    assert abs(r["DEr"].sum() + r["DEt"].sum() - 1.0) <= 1e-12


def test_planted_cx_cy_swap_fails(monkeypatch):
    case = [c for c in CASES if c["name"] == "lx_ne_ly"][0]
    orig = rcwa2d.layer_operators

    def swapped(*a, **k):
        e, cx, cy = orig(*a, **k)
        return e, cy, cx
    monkeypatch.setattr(rcwa2d, "layer_operators", swapped)
    j = case["jones"]
    r = rcwa2d.solve(np.array(case["stack"]), case["wavelength"], case["theta"], case["phi"], case["M1"], case["M2"],
                     (complex(j[0], j[1]), complex(j[2], j[3])))
    assert np.max(np.abs(r["DEr"] - np.asarray(case["DEr"]))) > 1e-6
