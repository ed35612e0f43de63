"""GPU parity of the conical solve through the C-ABI.

* the reference's ten Jones-matrix tables, |dJ| <= 1e-9 max(1, |J|)
* the reference's compiled solver on seeded random instances, DE at 1e-10
* degenerate layers take the Schur-basis propagation (same vectors)
* phi = 0 against the planar GPU path; lossless flux unitarity
* seeded instances of the benchmark's conical generator (M 20..40) against the oracle
* each result bitwise the same alone and in a mixed batch
"""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


class _S:
    def __init__(self, f):
        self.f = np.asarray(f, dtype=np.float64)

    def flat(self):
        return self.f


def _run(cases, ctx):
    from stablercwa_amd import rcwa
    b = rcwa.ConicalBatch([_S(c["stack"]) for c in cases],
                          [(i, c["wavelength"], c["theta"], c["phi"], c["M"]) for i, c in enumerate(cases)], ctx)
    b.execute()
    return b, b.results()


def _jerr(J, ref_flat):
    R = np.asarray(ref_flat).view(np.complex128).reshape(-1, 4)
    return float(np.max(np.abs(J - R) / np.maximum(1.0, np.abs(R))))


def test_conical_reference_tables(ctx):
    cases = load_golden("conical_tables.json")["cases"]
    b, res = _run(cases, ctx)
    for i, c in enumerate(cases):
        r = b.result(i, res)
        assert res["status"][i] == 0, c["id"]
        assert _jerr(r.Jr, c["Jr"]) <= 1e-9, c["id"]
        assert _jerr(r.Jt, c["Jt"]) <= 1e-9, c["id"]


def test_conical_reference_build(ctx):
    cases = [c for c in load_golden("conical_ref.json")["cases"] if c["n_schur"] == 0]
    b, res = _run(cases, ctx)
    for i, c in enumerate(cases):
        r = b.result(i, res)
        assert res["status"][i] == 0, c["name"]
        assert np.max(np.abs(r.DEr - np.reshape(c["DEr"], (-1, 4)))) <= 1e-10, c["name"]
        assert np.max(np.abs(r.DEt - np.reshape(c["DEt"], (-1, 4)))) <= 1e-10, c["name"]


def test_conical_schur_layers(ctx):
    cases = [c for c in load_golden("conical_ref.json")["cases"] if c["n_schur"] > 0]
    assert cases
    b, res = _run(cases, ctx)
    for i, c in enumerate(cases):
        r = b.result(i, res)
        assert res["status"][i] == 0, c["name"]
        assert np.max(np.abs(r.DEr - np.reshape(c["DEr"], (-1, 4)))) <= 1e-10, c["name"]
        assert np.max(np.abs(r.DEt - np.reshape(c["DEt"], (-1, 4)))) <= 1e-10, c["name"]


def test_phi_zero_matches_planar_gpu(ctx):
    from stablercwa_amd import rcwa
    from stablercwa_amd.workloads import mirror_stack
    st = mirror_stack()
    c = rcwa.solve_conical(st, 1.02, 40.0, 0.0, 15, ctx)
    te = rcwa.solve(st, 1.02, 40.0, 15, rcwa.Polarization.TE, ctx)
    tm = rcwa.solve(st, 1.02, 40.0, 15, rcwa.Polarization.TM, ctx)
    assert np.max(np.abs(c.DEr[:, 0] - te.DEr)) <= 1e-11 and np.max(np.abs(c.DEt[:, 0] - te.DEt)) <= 1e-11
    assert np.max(np.abs(c.DEr[:, 3] - tm.DEr)) <= 1e-11 and np.max(np.abs(c.DEt[:, 3] - tm.DEt)) <= 1e-11


def test_generator_instances_vs_oracle(ctx):
    from oracle import conical
    from stablercwa_amd import rcwa
    from stablercwa_amd.workloads import conical_batch
    probs = conical_batch(16, seed=0xC3, m_lo=20, m_hi=40)
    b = rcwa.ConicalBatch([p.stack for p in probs], [(i, p.wavelength, p.theta, p.phi, p.M) for i, p in enumerate(probs)], ctx)
    b.execute()
    res = b.results()
    for i, p in enumerate(probs):
        o = conical.solve(p.stack.flat(), p.wavelength, p.theta, p.phi, p.M)
        r = b.result(i, res)
        assert res["status"][i] == 0
        assert np.max(np.abs(r.DEr - o["DEr"])) <= 1e-9, i
        assert np.max(np.abs(r.DEt - o["DEt"])) <= 1e-9, i
        if i % 2 == 0:     # lossless instances: both input polarisations conserve energy
            for col in ((0, 2), (1, 3)):
                tot = r.DEr[:, list(col)].sum() + r.DEt[:, list(col)].sum()
                assert abs(tot - 1.0) <= 1e-11


def test_conical_alone_equals_in_mix(ctx):
    cases = load_golden("conical_ref.json")["cases"][:10]
    b, mixed = _run(cases, ctx)
    for i in (0, 4, 9):
        b1, alone = _run([cases[i]], ctx)
        assert np.array_equal(b.result(i, mixed).Jr, b1.result(0, alone).Jr)
        assert np.array_equal(b.result(i, mixed).Jt, b1.result(0, alone).Jt)


def test_conical_invalid_input(ctx):
    from stablercwa_amd import rcwa
    from stablercwa_amd.types import Layer, Stack
    st = Stack(1.0, [Layer.grating(0.3, [0, 0.2, 0.6], [1.5, 1.0])], complex(1.5, 0.1), 0.6)
    with pytest.raises(ValueError):
        rcwa.solve_conical(st, 1.0, 10.0, 20.0, 5, ctx)       # lossy exit medium
    st = Stack(1.0, [Layer.grating(0.3, [0, 0.2, 0.6], [1.5, 1.0])], 1.5, 0.6)
    with pytest.raises(ValueError):
        rcwa.solve_conical(st, 1.0, 10.0, float("nan"), 5, ctx)


def test_conical_geometry_and_flux_unitarity(ctx):
    """Per-order kx, kz, flux factors and phi_m against the oracle; flux unitarity is the identity when lossless."""
    from oracle import conical
    cases = load_golden("conical_tables.json")["cases"]
    b, res = _run(cases, ctx)
    for i, c in enumerate(cases):
        r = b.result(i, res)
        o = conical.solve(np.array(c["stack"]), c["wavelength"], c["theta"], c["phi"], c["M"])
        assert np.max(np.abs(r.flux_r - o["flux_r"])) <= 1e-14, c["id"]
        assert np.max(np.abs(r.flux_t - o["flux_t"])) <= 1e-14, c["id"]
        nI = float(c["stack"][1])
        sine = nI * np.sin(np.radians(c["theta"]))
        kx = sine * np.cos(np.radians(c["phi"])) - np.arange(-c["M"], c["M"] + 1) * c["wavelength"] / c["stack"][0]
        assert np.max(np.abs(r.kx - kx)) <= 1e-15 * max(1.0, np.max(np.abs(kx))), c["id"]
        assert np.all(r.kzI.imag >= 0.0) and np.all(np.abs(r.phi_m_deg) <= 180.0), c["id"]
    r = b.result(0, res)                                   # binary grating, lossless
    assert np.max(np.abs(r.flux_unitarity() - np.eye(2))) <= 1e-12
