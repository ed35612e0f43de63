"""The JSON request/reply layer (include/stablercwa_amd/rcwa_json.hpp), the parts that need no GPU: the parser
and writer through the command-line driver, and the request checks (the reply is {"ok": false, "error", "where"}
with the field path of the first problem, as the reference's tests/test_rcwa_json.cpp:1058-1237 and
test_rcwa2d_json.cpp require; messages are this build's own).  Every case edits one part of a good request.
The solved replies are checked against direct solver calls in tests/test_gpu_json.py.
"""
import json
import os
import subprocess

import pytest

from conftest import ROOT

DRIVER = os.path.join(ROOT, "stablercwa_amd", "_obj", "json_driver")

LAYERS = ('[{"d": 0.38, "edges": [0.0, 0.1722, 0.574], "n_values": [1.46, 1.0]}, '
          '{"d": 0.115, "n": [1.96, 0.0]}]')
PLANAR = {
    "mode": '"mode": "planar"',
    "stack": '"stack": {"nI": 1.0, "nII": 1.46, "period": 0.574, "layers": ' + LAYERS + '}',
    "M": '"M": 20',
    "pol": '"polarization": "TE"',
    "points": '"points": [{"wavelength": 1.053, "theta_deg": 66.36}, {"wavelength": 1.1, "theta_deg": 10.0}]',
    "options": '"options": {"diagnostics": false}',
    "extra": "",
}
CONICAL = dict(PLANAR, mode='"mode": "conical"', pol="",
               points='"points": [{"wavelength": 1.053, "theta_deg": 40.0, "phi_deg": 30.0}]',
               extra='"input": {"psi_deg": 45.0, "delta_deg": 90.0}')
ROWS = ('[{"edges": [0.0, 0.8], "n_values": [1.0]}, {"edges": [0.0, 0.25, 0.55, 0.8], "n_values": [1.0, 1.8, 1.0]}, '
        '{"edges": [0.0, 0.8], "n_values": [1.0]}]')
PERIODIC = {
    "mode": '"mode": "periodic2d"',
    "stack": ('"stack": {"nI": 1.0, "nII": 1.46, "Lx": 0.8, "Ly": 0.8, "layers": [{"d": 0.3, "y_edges": '
              '[0.0, 0.25, 0.55, 0.8], "x_profiles": ' + ROWS + '}, {"d": 0.115, "n": [1.96, 0.0]}]}'),
    "M": '"M": 2, "My": 2',
    "pol": "",
    "points": '"points": [{"wavelength": 1.053, "theta_deg": 42.0, "phi_deg": -25.0}]',
    "options": '"options": {"factorization": "li"}',
    "extra": '"input": {"psi_deg": 90.0, "delta_deg": 0.0}',
}


def text(parts):
    return "{" + ", ".join(v for v in parts.values() if v) + "}"


def edit(base, part, old, new):
    parts = dict(base)
    if old is None:
        parts[part] = new
    else:
        assert parts[part].count(old) == 1, (part, old)
        parts[part] = parts[part].replace(old, new)
    return text(parts)


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        pytest.skip("json_driver is not built (python -m stablercwa_amd.build)")
    return DRIVER


def reply(driver, request, indent=None):
    args = [driver, "-"] + (["--indent", str(indent)] if indent is not None else [])
    r = subprocess.run(args, input=request.encode(), capture_output=True, timeout=60)
    return r.returncode, r.stdout.decode()


def rejected_at(driver, request, where):
    rc, out = reply(driver, request)
    v = json.loads(out)
    assert rc == 1 and list(v) == ["ok", "error", "where"] and v["ok"] is False and v["error"], out
    assert v["where"] == where, (where, out)


PLANAR_BAD = [
    ("mode", None, "", "mode"),
    ("mode", '"planar"', "3", "mode"),
    ("mode", '"planar"', '"Planar"', "mode"),
    ("extra", None, '"polarisation": "TE"', "polarisation"),
    ("extra", None, '"phi_deg": 10', "phi_deg"),
    ("extra", None, '"input": {"psi_deg": 90}', "input"),
    ("stack", None, "", "stack"),
    ("stack", None, '"stack": [1, 2]', "stack"),
    ("stack", '"period": 0.574', '"period": 0.574, "nIII": 1', "stack.nIII"),
    ("stack", '"nI": 1.0, ', "", "stack.nI"),
    ("stack", '"nI": 1.0', '"nI": "air"', "stack.nI"),
    ("stack", '"nI": 1.0', '"nI": [1.0, 0.0, 0.0]', "stack.nI"),
    ("stack", '"nI": 1.0', '"nI": [1.0, 0.1]', "stack.nI"),
    ("stack", '"nI": 1.0', '"nI": -1.0', "stack.nI"),
    ("stack", '"nII": 1.46', '"nII": [1.5, -0.1]', "stack.nII"),
    ("stack", '"nII": 1.46', '"nII": true', "stack.nII"),
    ("stack", '"period": 0.574', '"period": 0', "stack.period"),
    ("stack", '"period": 0.574', '"period": 0.6', "stack.layers[0].edges"),
    ("stack", ', "layers": ' + LAYERS, "", "stack.layers"),
    ("stack", LAYERS, "{}", "stack.layers"),
    ("stack", LAYERS, "[3]", "stack.layers[0]"),
    ("stack", '"d": 0.115', '"d": 0.115, "thickness": 1', "stack.layers[1].thickness"),
    ("stack", '"d": 0.38, ', "", "stack.layers[0].d"),
    ("stack", '"d": 0.38', '"d": -0.38', "stack.layers[0].d"),
    ("stack", '"n": [1.96, 0.0]', '"n": [1.96, 0.0], "edges": [0.0, 0.574]', "stack.layers[1]"),
    ("stack", ', "n": [1.96, 0.0]', "", "stack.layers[1]"),
    ("stack", ', "n_values": [1.46, 1.0]', "", "stack.layers[0].n_values"),
    ("stack", '"edges": [0.0, 0.1722, 0.574], ', "", "stack.layers[0].edges"),
    ("stack", "[0.0, 0.1722, 0.574]", '[0.0, "a", 0.574]', "stack.layers[0].edges[1]"),
    ("stack", "[0.0, 0.1722, 0.574]", "[0.0, 0.3, 0.2]", "stack.layers[0].edges[2]"),
    ("stack", "[1.46, 1.0]", "[1.46, [1.0]]", "stack.layers[0].n_values[1]"),
    ("stack", "[1.46, 1.0]", "[0.0, 1.0]", "stack.layers[0].n_values[0]"),
    ("stack", "[1.46, 1.0]", "[1.46, [1.0, -0.2]]", "stack.layers[0].n_values[1]"),
    ("stack", "[1.46, 1.0]", "[1.46, 1.0, 2.0]", "stack.layers[0].n_values"),
    ("stack", '"n_values": [1.46, 1.0]', '"n_values": [1.46, 1.0], "fourier_method": "dft"',
     "stack.layers[0].fourier_method"),
    ("stack", '"n_values": [1.46, 1.0]', '"n_values": [1.46, 1.0], "fft_samples": -1', "stack.layers[0].fft_samples"),
    ("M", None, "", "M"),
    ("M", "20", "-1", "M"),
    ("M", "20", "20.5", "M"),
    ("M", "20", '"20"', "M"),
    ("M", "20", "8192", "M"),
    ("pol", None, "", "polarization"),
    ("pol", '"TE"', '"te"', "polarization"),
    ("points", None, "", "points"),
    ("points", None, '"points": []', "points"),
    ("points", None, '"points": {}', "points"),
    ("points", None, '"points": [{"wavelength": 1.0, "theta_deg": 1.0}, 5]', "points[1]"),
    ("points", '"wavelength": 1.1, ', "", "points[1].wavelength"),
    ("points", '"wavelength": 1.1', '"wavelength": -1.1', "points[1].wavelength"),
    ("points", '"theta_deg": 10.0', '"theta_deg": 90.0', "points[1].theta_deg"),
    ("points", '"theta_deg": 10.0', '"theta_deg": 10.0, "phi_deg": 5.0', "points[1].phi_deg"),
    ("points", '"theta_deg": 10.0', '"theta_deg": 10.0, "lambda": 5.0', "points[1].lambda"),
    ("points", None, '"points": [{"wavelength": 1.0, "theta_deg": 1.0}], "wavelength": 1.0', "points"),
    ("options", None, '"options": 3', "options"),
    ("options", '"diagnostics": false', '"diagnostics": 0', "options.diagnostics"),
    ("options", '"diagnostics": false', '"speed": 1', "options.speed"),
    ("options", '"diagnostics": false', '"eig": "lr"', "options.eig"),
    ("options", '"diagnostics": false', '"eig": "jacobi"', "options.eig"),
    ("options", '"diagnostics": false', '"linear": "qr"', "options.linear"),
    ("options", '"diagnostics": false', '"backend": "mkl"', "options.backend"),
    ("options", '"diagnostics": false', '"homogeneous_fast_path": false', "options.homogeneous_fast_path"),
    ("options", '"diagnostics": false', '"threads": 0', "options.threads"),
    ("options", '"diagnostics": false', '"threads": 257', "options.threads"),
    ("options", '"diagnostics": false', '"factorization": "li"', "options.factorization"),
]

CONICAL_BAD = [
    ("extra", None, '"polarization": "TE"', "polarization"),
    ("extra", None, '"input": {"psi_deg": 45.0, "jones": [1, 0]}', "input"),
    ("extra", None, '"input": {"jones": [1]}', "input.jones"),
    ("extra", None, '"input": {"jones": [0, [0, 0]]}', "input.jones"),
    ("extra", None, '"input": {"jones": [1, "x"]}', "input.jones[1]"),
    ("extra", None, '"input": {"delta_deg": 3}', "input.psi_deg"),
    ("extra", None, '"input": {"psi_deg": "45"}', "input.psi_deg"),
    ("extra", None, '"input": {"psi_deg": 45, "chi": 1}', "input.chi"),
    ("stack", '"nII": 1.46', '"nII": [1.46, 0.01]', "stack.nII"),
    ("points", '"phi_deg": 30.0', '"phi_deg": "30"', "points[0].phi_deg"),
    ("M", "20", "4096", "M"),
]

PERIODIC_BAD = [
    ("M", None, '"M": 2', "My"),
    ("M", '"My": 2', '"My": -1', "My"),
    ("M", '"My": 2', '"My": 300', "My"),
    ("extra", None, "", "input"),
    ("extra", None, '"input": {"psi_deg": 90.0}, "polarization": "TE"', "polarization"),
    ("stack", '"Lx": 0.8', '"Lx": 0.8, "period": 0.8', "stack.period"),
    ("stack", '"Lx": 0.8, ', "", "stack.Lx"),
    ("stack", '"Ly": 0.8', '"Ly": -0.8', "stack.Ly"),
    ("stack", '"y_edges": [0.0, 0.25, 0.55, 0.8]', '"y_edges": [0.0, 0.25, 0.55, 0.9]',
     "stack.layers[0].y_edges"),
    ("stack", '"y_edges": [0.0, 0.25, 0.55, 0.8]', '"y_edges": [0.0, 0.55, 0.25, 0.8]',
     "stack.layers[0].y_edges[2]"),
    ("stack", '"y_edges": [0.0, 0.25, 0.55, 0.8]', '"y_edges": [0.0, 0.8]', "stack.layers[0].x_profiles"),
    ("stack", '"n_values": [1.0, 1.8, 1.0]', '"n_values": [1.0, 1.8]', "stack.layers[0].x_profiles[1].n_values"),
    ("stack", '"n_values": [1.0, 1.8, 1.0]', '"n_values": [1.0, [1.8, -1], 1.0]',
     "stack.layers[0].x_profiles[1].n_values[1]"),
    ("stack", '[0.0, 0.25, 0.55, 0.8], "n_values"', '[0.0, 0.25, 0.55, 0.7], "n_values"',
     "stack.layers[0].x_profiles[1].edges"),
    ("stack", '"d": 0.3', '"d": 0.3, "edges": [0, 1]', "stack.layers[0].edges"),
    ("options", '"factorization": "li"', '"factorization": "fft"', "options.factorization"),
    ("options", '"factorization": "li"', '"linear": "lu"', "options.linear"),
]


@pytest.mark.parametrize("case", PLANAR_BAD, ids=lambda c: c[3] or "top")
def test_planar_request_errors(driver, case):
    part, old, new, where = case
    rejected_at(driver, edit(PLANAR, part, old, new), where)


@pytest.mark.parametrize("case", CONICAL_BAD, ids=lambda c: c[3])
def test_conical_request_errors(driver, case):
    part, old, new, where = case
    rejected_at(driver, edit(CONICAL, part, old, new), where)


@pytest.mark.parametrize("case", PERIODIC_BAD, ids=lambda c: c[3])
def test_periodic2d_request_errors(driver, case):
    part, old, new, where = case
    rejected_at(driver, edit(PERIODIC, part, old, new), where)


@pytest.mark.parametrize("bad", ['', '   ', '{"mode": "planar"', '[1, 2]', '{"a": 1, "a": 2}', '{"a": tru}',
                                 '{"a": 01}', '{"a": 1.}', '{"a": "\\x"}', '{"a": "\\ud800"}', '{"a": 1e400}',
                                 '{"a": 1} x', '{"a": "\x01"}'])
def test_syntax_errors_have_empty_where(driver, bad):
    rc, out = reply(driver, bad)
    v = json.loads(out)
    assert rc == 1 and v["ok"] is False and v["where"] == "", out


def test_syntax_error_position(driver):
    rc, out = reply(driver, '{\n  "mode": "planar",\n  "M": 2x\n}')
    v = json.loads(out)
    assert "line 3, column 9" in v["error"], v


def test_good_requests_pass_the_checks(driver):
    """The unedited requests get past every check: what comes back is a solved reply (on a GPU box) or the
    solver's own error with an empty path (here, where there is no GPU), never a request error."""
    for base in (PLANAR, CONICAL, PERIODIC):
        rc, out = reply(driver, text(base))
        v = json.loads(out)
        assert v["ok"] is True or (v["where"] == "" and v["error"].startswith("solver error")), out


def test_indent_and_escapes(driver):
    req = edit(PLANAR, "extra", None, '"bad\\n\\"key": 1')
    rc, out = reply(driver, req, indent=2)
    assert out.startswith('{\n  "ok": false,\n  "error": ')
    v = json.loads(out)
    assert v["where"] == 'bad\n"key'
