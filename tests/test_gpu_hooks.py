"""GPU: the hook adapters (include/stablercwa_amd/hooks.hpp) bound into the reference's own solver.

oracle/_ref/ref_hooks is the reference's header-only solver built from its headers, with
Options::custom_eig / custom_solve / custom_multiply (rcwa1d.hpp:122-136; Options2D, rcwa2d_types.hpp:89-95)
set to this build's GPU adapters, exactly as a maintainer of a reference caller would bind them.  Each
hooked solve is compared with the same binary's built-in solve:

* planar TE/TM, conical and crossed solves with each hook and with all three: efficiencies agree to
  1e-10, the solve makes the same numbers of eigensolves and dense solves, and it really went through
  the bound hooks (the driver counts their calls)
* custom_eig with max_iter = 1: the adapter returns info > 0 and no vectors, and the reference's solver
  reports eig_info > 0 with NaN results (its own idiom, tests/test_rcwa1d.cpp:1169-1181)
* custom_solve on a singular interface (an order grazing exactly in the exit medium): the adapter throws
  std::runtime_error and the solver returns NaN results as with its built-in LU (rcwa1d.hpp:272-275)
* the adapters' own contracts: unit 2-norm eigenvectors with a residual at the reference's level,
  solve and product equal to the reference's stablercwa::solve and operator*, std::invalid_argument
  for wrong shapes, std::runtime_error for an exact zero pivot
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_hooks")
MODES = ("eig", "solve", "multiply", "all")


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        pytest.skip("oracle/_ref/ref_hooks not built (the reference tree was absent at build time)")
    return DRIVER


def _run(driver, lines):
    r = subprocess.run([driver], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    blocks, cur = [], []
    for ln in r.stdout.splitlines():
        if ln == "end":
            blocks.append(cur)
            cur = []
        else:
            cur.append(ln)
    assert len(blocks) == len(lines)
    out = []
    for b in blocks:
        head = b[0].split()
        f = {"head": head}
        for ln in b[1:]:
            parts = ln.split()
            f[parts[0]] = np.array([float(x) for x in parts[1:]]) if head[0] == "ok" else parts[1:]
        out.append(f)
    return out


def _rec(flat):
    return f"{len(flat)} " + " ".join(repr(float(x)) for x in flat)


def _grating(nII=1.45):
    from stablercwa_amd.types import Layer, Stack
    p, f = 0.5747, 0.30
    return Stack(1.0, [Layer.grating(0.38, [-f * p / 2, f * p / 2, p - f * p / 2], [1.45, 1.0])], nII, p)


def _planar_cases():
    from stablercwa_amd.types import Layer, Stack
    from stablercwa_amd.workloads import mirror_stack
    two = Stack(1.2, [Layer.grating(0.2, [0.0, 0.2, 0.5], [1.8 + 0.1j, 1.0]), Layer.uniform(0.1, 2.0 + 0.05j),
                      Layer.grating(0.3, [0.0, 0.25, 0.4, 0.5], [1.5, 2.2, 1.0])], 1.5 + 0.03j, 0.5)
    return [(_grating(), 1.053, 66.36, 6, 0), (_grating(), 1.053, 66.36, 6, 1), (two, 0.9, 20.0, 8, 1),
            (mirror_stack(), 1.0, 66.36, 15, 0)]


HOOK = {"eig": (0,), "solve": (1,), "multiply": (2,), "all": (0, 1, 2)}


def _same(a, b, what, mode, n_eig_field=True):
    assert a["head"][0] == b["head"][0] == "ok", (what, a["head"], b["head"])
    assert a["head"][1:] == b["head"][1:], (what, a["head"], b["head"])       # eig_info and the counts
    err = max(np.max(np.abs(a["DEr"] - b["DEr"])), np.max(np.abs(a["DEt"] - b["DEt"])))
    assert err <= 1e-10, (what, err)
    # the solver went through the bound hooks (custom_eig once per eigensolve), the built-in through none
    assert np.all(b["calls"] == 0), (what, b["calls"])
    for k in HOOK[mode]:
        assert a["calls"][k] > 0, (what, mode, a["calls"])
    if 0 in HOOK[mode] and n_eig_field:
        # planar and conical heads carry n_eig in field 2 (conical: n_schur in field 4, whose extra 2N Schur
        # eigensolve always takes the reference's built-in QR, rcwa1d_conical.hpp:63)
        n_schur = int(b["head"][4]) if len(b["head"]) > 4 else 0
        assert a["calls"][0] == int(b["head"][2]) - n_schur, (what, a["calls"], b["head"])


def test_planar_hooks_equal_builtin(driver):
    lines = []
    for st, lam, th, M, pol in _planar_cases():
        for mode in ("builtin",) + MODES:
            lines.append(f"planar {mode} {pol} {M} {lam!r} {th!r} {_rec(st.flat())}")
    res = _run(driver, lines)
    k = len(MODES) + 1
    for i in range(len(_planar_cases())):
        base = res[k * i]
        for j, mode in enumerate(MODES):
            _same(res[k * i + 1 + j], base, (i, mode), mode)


def test_conical_and_crossed_hooks_equal_builtin(driver):
    from stablercwa_amd.types import Stack
    from stablercwa_amd.workloads import pillar_batch
    lines = []
    c = load_golden("conical_ref.json")["cases"]
    conical = [c[0], [x for x in c if x["name"].startswith("schur")][0]]
    for q in conical:
        st = Stack.from_flat(q["stack"])
        for mode in ("builtin",) + MODES:
            lines.append(f"conical {mode} {q['M']} {q['wavelength']!r} {q['theta']!r} {q['phi']!r} {_rec(st.flat())}")
    p = pillar_batch(1, seed=0xD1C, orders=(2,))[0]
    for mode in ("builtin",) + MODES:
        lines.append(f"crossed {mode} 2 2 {p.wavelength!r} {p.theta!r} {p.phi!r} 0.6 0.1 0.78 0.0 {_rec(p.stack.flat())}")
    res = _run(driver, lines)
    k = len(MODES) + 1
    for i in range(len(conical) + 1):
        for j, mode in enumerate(MODES):
            _same(res[k * i + 1 + j], res[k * i], (i, mode), mode, n_eig_field=i < len(conical))


def test_unconverged_gpu_eig_gives_nan(driver):
    """The same result through the GPU adapter with max_iter = 1 as through the reference's own eig_qr with
    max_iter = 1 (its test idiom): eig_info > 0 (the same index) and NaN results."""
    st = _grating()
    gpu, cpu = _run(driver, [f"planar {m} 0 6 1.053 66.36 {_rec(st.flat())}" for m in ("eig1", "ref1")])
    for r in (gpu, cpu):
        assert r["head"][0] == "ok" and int(r["head"][1]) > 0 and r["calls"][0] == 1
        assert np.all(np.isnan(r["DEr"])) and np.all(np.isnan(r["DEt"]))
    assert gpu["head"][1] == cpu["head"][1]


def test_singular_gpu_solve_gives_nan(driver):
    st = _grating(nII=2.0)             # orders +-2 graze exactly in the exit medium at lambda = period
    lines = [f"planar {mode} 0 8 {st.period!r} 0.0 {_rec(st.flat())}" for mode in ("builtin", "solve")]
    builtin, hooked = _run(driver, lines)
    for r in (builtin, hooked):
        assert r["head"][0] == "ok"
        assert np.all(np.isnan(r["DEr"])) and np.all(np.isnan(r["DEt"]))
    assert hooked["calls"][1] > 0


def test_adapter_contracts(driver):
    b = _run(driver, ["contract"])[0]
    vals = {}
    for ln in [" ".join(b["head"])] + [" ".join([k] + list(v)) for k, v in b.items() if k != "head"]:
        parts = ln.split()
        vals[parts[0]] = parts[1:]
    assert vals["eig_info"] == ["0"]
    assert float(vals["eig_residual"][0]) <= 2e-15
    assert float(vals["eig_unit_norm_defect"][0]) <= 1e-14
    assert int(vals["eig1_info"][0]) > 0 and vals["eig1_vectors_empty"] == ["1"]
    assert float(vals["solve_max_diff"][0]) <= 1e-12
    assert float(vals["solve_backward_error"][0]) <= 1e-15
    assert float(vals["multiply_max_diff"][0]) <= 1e-12
    assert vals["multiply_shape"] == ["37", "5"]
    assert vals["eig_nonsquare"] == ["invalid_argument"]
    assert vals["multiply_mismatch"] == ["invalid_argument"]
    assert vals["solve_mismatch"] == ["invalid_argument"]
    assert vals["solve_singular"] == ["runtime_error"]
