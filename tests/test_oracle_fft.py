"""CPU: the oracle's FFT-sampled Fourier coefficients (oracle/rcwa1d.py, restating fourier.hpp:26-83), pinned by

* an independent O(N hmax) quadrature with every phase taken at the absolute sample coordinate (the
  reference's own check, tests/test_fft.cpp:19-43), at 3e-12, for several origins and sample counts
  including non-powers of two; it fixes the sign, the 1/N scaling, the bins and the midpoint/origin phase
* a discontinuity exactly at a midpoint belongs to the segment on its right, and a constant profile gives
  its value (tests/test_fft.cpp:45-48)
* convergence to the analytic coefficients as N grows (tests/test_fft.cpp:50-59)
* a planted mistake (the half-cell phase dropped) that the quadrature catches
* whole solves with FFT layers against the reference's compiled solver (tests/golden/fft_ref.json) at 1e-10
* the reference's sample-count rules and messages (resolve_fft_samples, fourier.hpp:26-44)
"""
import math

import numpy as np
import pytest

from conftest import load_golden


def direct_dft(edges, values, period, hmax, N):
    out = np.zeros(2 * hmax + 1, dtype=complex)
    for j in range(N):
        u = (j + 0.5) / N * period
        k = 0
        while k + 1 < len(values) and u >= edges[k + 1] - edges[0]:
            k += 1
        for h in range(-hmax, hmax + 1):
            a = -2 * math.pi * h * (edges[0] + u) / period
            out[h + hmax] += values[k] * complex(math.cos(a), math.sin(a)) / N
    return out


def test_fft_coefficients_match_direct_quadrature():
    from oracle.rcwa1d import fourier_coefficients_fft
    values = [2.3 + 0.4j, 1.1 - 0.2j, 3.0 + 0.1j]
    for origin in (0.0, -0.3, 0.2, 1.2):
        edges = [origin, origin + 0.3125, origin + 0.7, origin + 1.0]
        for N in (9, 17, 60, 257, 8192):
            got = fourier_coefficients_fft(edges, values, 1.0, 4, N)
            want = direct_dft(edges, values, 1.0, 4, N)
            assert np.max(np.abs(got - want)) < 3e-12, (origin, N)


def test_fft_boundary_and_constant():
    from oracle.rcwa1d import fourier_coefficients_fft
    b = fourier_coefficients_fft([0.0, 0.5 / 9, 1.0], [4.0, 1.0], 1.0, 4, 9)
    assert abs(b[4] - 1.0) < 1e-14
    c = fourier_coefficients_fft([0.0, 1.0], [2.0 + 0.3j], 1.0, 0, 1)
    assert abs(c[0] - (2.0 + 0.3j)) < 1e-14


def test_fft_converges_to_analytic():
    from oracle.rcwa1d import fourier_coefficients, fourier_coefficients_fft
    edges, values = [-0.3, -0.05, 0.7], [2.0 + 0.3j, 1.0 - 0.1j]
    exact = fourier_coefficients(edges, values, 1.0, 4)
    previous = 1.0
    for N in (64, 256, 1024):
        err = np.max(np.abs(fourier_coefficients_fft(edges, values, 1.0, 4, N) - exact))
        assert err < previous / 10
        previous = err


def test_planted_phase_mistake_is_caught(monkeypatch):
    """Dropping the half-cell phase (offset = origin / period) must fail the quadrature check."""
    import oracle.rcwa1d as o
    real = o.fourier_coefficients_fft

    def no_half_cell(edges, values, period, hmax, samples):
        c = real(edges, values, period, hmax, samples)
        h = np.arange(-hmax, hmax + 1)
        return c * np.exp(2j * math.pi * h * 0.5 / samples)     # undo the half-cell factor
    edges, values = [0.0, 0.3125, 0.7, 1.0], [2.3 + 0.4j, 1.1 - 0.2j, 3.0 + 0.1j]
    err = np.max(np.abs(no_half_cell(edges, values, 1.0, 4, 17) - direct_dft(edges, values, 1.0, 4, 17)))
    assert err > 1e-3


def test_oracle_fft_solves_match_reference():
    from oracle import conical, rcwa1d
    for c in load_golden("fft_ref.json")["cases"]:
        if c["kind"] == "planar":
            if c["M"] > 80:
                continue                       # the M = 130 case is for the GPU test (slow in numpy)
            o = rcwa1d.solve(c["stack"], c["wavelength"], c["theta"], c["M"], c["pol"])
            err = max(np.max(np.abs(o["DEr"] - c["DEr"])), np.max(np.abs(o["DEt"] - c["DEt"])))
        else:
            o = conical.solve(c["stack"], c["wavelength"], c["theta"], c["phi"], c["M"])
            err = max(np.max(np.abs(np.ravel(o["DEr"]) - c["DEr"])), np.max(np.abs(np.ravel(o["DEt"]) - c["DEt"])))
        assert err <= 1e-10, (c["name"], err)


def test_resolve_fft_samples_rules():
    from oracle.rcwa1d import resolve_fft_samples
    assert resolve_fft_samples(6) == 8192
    assert resolve_fft_samples(70) == 16384
    assert resolve_fft_samples(130) == 32768
    assert resolve_fft_samples(2, 9) == 9
    with pytest.raises(ValueError, match="at least 4\\*M\\+1"):
        resolve_fft_samples(2, 8)
    with pytest.raises(ValueError, match="automatic FFT exceeds"):
        resolve_fft_samples(40000)
