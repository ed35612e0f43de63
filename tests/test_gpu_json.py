"""GPU: solved replies of the JSON layer (include/stablercwa_amd/rcwa_json.hpp, through tests/cpp/json_driver.cpp).

Every number in a reply reads back bitwise as the solver returned it (the reference's contract,
rcwa_json.hpp:6-7): each reply is compared with the same solves made directly through the library, and the
README grating's orders with the reference's published values (README.md:106-111).  A numerical failure is
not a request error: ok stays true and the failed numbers are null, eig_info / status as the solver gives
them (planar exit-side grazing, tests/test_rcwa1d.cpp:1229-1232; a crossed-grating Rayleigh anomaly,
rcwa2d.hpp:1100-1115).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "stablercwa_amd", "_obj", "json_driver")


def run(request: dict):
    r = subprocess.run([DRIVER, "-"], input=json.dumps(request).encode(), capture_output=True, timeout=120)
    return r.returncode, json.loads(r.stdout.decode())


def cplx(a):
    return np.array([complex(x[0], x[1]) for x in a])


README = {"nI": 1.0, "nII": 1.46, "period": 0.574,
          "layers": [{"d": 0.38, "edges": [0.0, 0.172, 0.574], "n_values": [1.46, 1.0]}, {"d": 0.115, "n": [1.96, 0.0]}]}


def readme_stack():
    from stablercwa_amd.types import Layer, Stack
    return Stack(1.0, [Layer.grating(0.38, [0.0, 0.172, 0.574], [1.46, 1.0]), Layer.uniform(0.115, 1.96)], 1.46, 0.574)


def test_planar_reply_is_the_solver_bitwise(ctx):
    from stablercwa_amd import rcwa
    pts = [(1.053, 66.36), (1.0, 10.0), (1.2, 0.0)]
    for pol in ("TE", "TM"):
        rc, v = run({"mode": "planar", "stack": README, "M": 20, "polarization": pol,
                     "points": [{"wavelength": l, "theta_deg": t} for l, t in pts]})
        assert rc == 0 and v["ok"] is True and v["mode"] == "planar" and v["orders"] == list(range(-20, 21))
        assert list(v) == ["ok", "version", "mode", "M", "orders", "threads", "results"]
        b = rcwa.PlanarBatch([readme_stack()], [(0, l, t, 20, 0 if pol == "TE" else 1) for l, t in pts], ctx)
        b.execute()
        res = b.results()
        for k, r in enumerate(v["results"]):
            d = b.result(k, res)
            assert list(r)[:10] == ["wavelength", "theta_deg", "R", "T", "DEr", "DEt", "R_total", "T_total", "eig_info",
                                    "n_eig"]
            assert np.array_equal(np.array(r["DEr"]), d.DEr) and np.array_equal(np.array(r["DEt"]), d.DEt)
            assert np.array_equal(cplx(r["R"]), d.R) and np.array_equal(cplx(r["T"]), d.T)
            assert r["R_total"] == float(np.sum(d.DEr)) or abs(r["R_total"] - np.sum(d.DEr)) <= 1e-15
        b.close()
    # README.md:106-111: the README grating at M = 20, TE
    rc, v = run({"mode": "planar", "stack": README, "M": 20, "polarization": "TE", "wavelength": 1.053, "theta_deg": 66.36})
    r = v["results"][0]
    assert abs(r["DEr"][20] - 0.026034) < 5e-7 and abs(r["DEt"][20] - 0.497012) < 5e-7
    assert abs(r["DEr"][21] - 0.258135) < 5e-7 and abs(r["DEt"][21] - 0.218818) < 5e-7
    assert abs(r["R_total"] + r["T_total"] - 1.0) <= 1e-12


def test_conical_reply_with_input(ctx):
    from stablercwa_amd import rcwa
    stack = dict(README, layers=README["layers"][:1])
    rc, v = run({"mode": "conical", "stack": stack, "M": 12,
                 "points": [{"wavelength": 1.053, "theta_deg": 40.0, "phi_deg": 30.0}, {"wavelength": 1.1, "theta_deg": 5.0}],
                 "input": {"psi_deg": 45.0, "delta_deg": 90.0}})
    assert rc == 0 and v["ok"] is True
    from stablercwa_amd.types import Layer, Stack
    st = Stack(1.0, [Layer.grating(0.38, [0.0, 0.172, 0.574], [1.46, 1.0])], 1.46, 0.574)
    b = rcwa.ConicalBatch([st], [(0, 1.053, 40.0, 30.0, 12), (0, 1.1, 5.0, 0.0, 12)], ctx)
    b.execute()
    res = b.results()
    for k, r in enumerate(v["results"]):
        d = b.result(k, res)
        jr = np.array([[complex(*z) for z in o] for o in r["Jr"]])
        assert np.array_equal(jr, d.Jr)
        assert np.array_equal(np.array(r["DEr_channels"]), d.DEr) and np.array_equal(np.array(r["DEt_channels"]), d.DEt)
        assert np.array_equal(np.array(r["kx"]), d.kx) and np.array_equal(np.array(r["flux_r"]), d.flux_r)
        p = r["polarized"]
        assert list(p)[0] == "jones_in" and abs(p["R_total"] + p["T_total"] - 1.0) <= 1e-10
        assert r["min_abs_lambda"] is None or r["min_abs_lambda"] > 0.0
        assert isinstance(r["degenerate_mode_warning"], bool) and r["n_schur"] == 0
    b.close()


def test_periodic2d_reply_and_rayleigh(ctx):
    from stablercwa_amd import rcwa
    from stablercwa_amd.types import Layer2D, Stack2D, XProfile2D
    rows = [{"edges": [0.0, 0.8], "n_values": [1.0]}, {"edges": [0.0, 0.25, 0.55, 0.8], "n_values": [1.0, 1.8, 1.0]},
            {"edges": [0.0, 0.8], "n_values": [1.0]}]
    req = {"mode": "periodic2d",
           "stack": {"nI": 1.0, "nII": 1.46, "Lx": 0.8, "Ly": 0.8,
                     "layers": [{"d": 0.3, "y_edges": [0.0, 0.25, 0.55, 0.8], "x_profiles": rows},
                                {"d": 0.115, "n": [1.96, 0.0]}]},
           "M": 2, "My": 2, "points": [{"wavelength": 1.053, "theta_deg": 42.0, "phi_deg": -25.0},
                                       {"wavelength": 0.8, "theta_deg": 0.0}],
           "input": {"jones": [1.0, 0.0]}}
    rc, v = run(req)
    assert rc == 0 and v["ok"] is True and v["orders"][0] == [-2, -2] and len(v["orders"]) == 25
    r0, r1 = v["results"]
    assert r0["status"] == "ok"
    # lambda = Lx at normal incidence: orders (+-1, 0) graze exactly in the incidence medium
    assert r1["status"] == "rayleigh_anomaly" and r1["DEr"] == [None] * 25 and r1["Rs"][0] == [None, None]
    assert len(r1["prop_r"]) == 25
    st = Stack2D(nI=1.0, nII=1.46, Lx=0.8, Ly=0.8,
                 layers=[Layer2D.patterned(0.3, [0.0, 0.25, 0.55, 0.8],
                                           [XProfile2D([0.0, 0.8], [1.0]),
                                            XProfile2D([0.0, 0.25, 0.55, 0.8], [1.0, 1.8, 1.0]),
                                            XProfile2D([0.0, 0.8], [1.0])]),
                         Layer2D.uniform(0.115, 1.96)])
    b = rcwa.CrossedBatch([st], [(0, 1.053, 42.0, -25.0, 2, 2, (1.0 + 0j, 0j))], ctx)
    b.execute()
    d = b.result(0, b.results())
    b.close()
    assert np.array_equal(np.array(r0["DEr"]), d.DEr) and np.array_equal(np.array(r0["DEt"]), d.DEt)
    assert np.array_equal(cplx(r0["Rs"]), d.Rs)


def test_planar_numerical_failure_is_null_not_error(ctx):
    p, f = 0.5747, 0.30
    stack = {"nI": 1.0, "nII": 2.0, "period": p,
             "layers": [{"d": 0.38, "edges": [-f * p / 2, f * p / 2, p - f * p / 2], "n_values": [1.45, 1.0]}]}
    rc, v = run({"mode": "planar", "stack": stack, "M": 8, "polarization": "TE", "wavelength": p, "theta_deg": 0.0})
    assert rc == 0 and v["ok"] is True
    r = v["results"][0]
    assert r["DEr"] == [None] * 17 and r["R"][0] == [None, None] and r["R_total"] is None
