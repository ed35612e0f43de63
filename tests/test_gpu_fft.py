"""GPU: grating layers with FFT-sampled Fourier coefficients (FourierMethod::FFT, fourier.hpp:26-83) through the
C-ABI: the device samples eps (and 1/eps independently) at the cell midpoints, transforms with its own
radix-2 Stockham FFT (power-of-two sample counts) or a direct DFT of the needed bins (other counts), and
applies the origin and half-cell phase.

* planar TE/TM and conical solves against the reference's compiled solver (tests/golden/fft_ref.json) at
  1e-10: automatic counts 8192, 16384 and 32768 (M = 6, 70, 130), explicit 9, 257, 1000 and 4096, and
  gratings with a negative or shifted origin
* random FFT stacks against the numpy oracle at 1e-10, with the energy of lossless ones
* the reference's validation messages for the FFT settings (fourier.hpp:17-44)
"""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def test_fft_layers_against_reference(ctx):
    from stablercwa_amd import rcwa
    from stablercwa_amd.types import Stack
    cases = load_golden("fft_ref.json")["cases"]
    planar = [c for c in cases if c["kind"] == "planar"]
    conical = [c for c in cases if c["kind"] == "conical"]
    assert {c["M"] for c in planar} >= {2, 6, 70, 130}
    b = rcwa.PlanarBatch([Stack.from_flat(c["stack"]) for c in planar],
                         [(i, c["wavelength"], c["theta"], c["M"], c["pol"]) for i, c in enumerate(planar)], ctx)
    b.execute()
    res = b.results()
    for i, c in enumerate(planar):
        r = b.result(i, res)
        assert res["status"][i] == 0, c["name"]
        err = max(np.max(np.abs(r.DEr - c["DEr"])), np.max(np.abs(r.DEt - c["DEt"])))
        assert err <= 1e-10, (c["name"], err)
    b.close()
    b = rcwa.ConicalBatch([Stack.from_flat(c["stack"]) for c in conical],
                          [(i, c["wavelength"], c["theta"], c["phi"], c["M"]) for i, c in enumerate(conical)], ctx)
    b.execute()
    res = b.results()
    for i, c in enumerate(conical):
        r = b.result(i, res)
        assert res["status"][i] == 0, c["name"]
        err = max(np.max(np.abs(r.DEr.ravel() - c["DEr"])), np.max(np.abs(r.DEt.ravel() - c["DEt"])))
        assert err <= 1e-10, (c["name"], err)
    b.close()


def test_random_fft_stacks_against_oracle(ctx):
    from oracle import rcwa1d
    from stablercwa_amd import rcwa
    from stablercwa_amd.types import FourierMethod, Layer, Stack
    rng = np.random.default_rng(91)
    stacks, probs = [], []
    for i in range(12):
        period = float(rng.uniform(0.4, 1.5))
        lossless = i % 2 == 0
        layers = []
        for _ in range(int(rng.integers(1, 4))):
            K = int(rng.integers(2, 5))
            origin = float(rng.uniform(-0.5, 0.5)) * period
            e = np.sort(rng.uniform(0, period, K - 1))
            edges = [origin] + [origin + float(v) for v in e] + [origin + period]
            vals = [complex(rng.uniform(1.0, 2.8), 0.0 if lossless else rng.uniform(0, 0.1)) for _ in range(K)]
            M = int(rng.integers(3, 12))
            samples = [0, 4 * M + 1, 1000, 2048][int(rng.integers(0, 4))]
            layers.append(Layer.grating(float(rng.uniform(0.05, 0.5)), edges, vals, FourierMethod.FFT, samples))
            if rng.random() < 0.5:
                layers.append(Layer.uniform(float(rng.uniform(0.05, 0.3)), 1.6))
        stacks.append(Stack(1.0, layers, 1.45, period))
        # one M per stack, large enough for every explicit sample count in it
        m_max = min((s.fft_samples - 1) // 4 for s in layers if not s.homogeneous and s.fft_samples) \
            if any(not s.homogeneous and s.fft_samples for s in layers) else 10
        probs.append((i, float(rng.uniform(0.6, 1.4)) * period, float(rng.uniform(0, 60)), int(min(m_max, 10)),
                      i % 2))
    b = rcwa.PlanarBatch(stacks, probs, ctx)
    b.execute()
    res = b.results()
    for i, (s, lam, th, M, pol) in enumerate(probs):
        r = b.result(i, res)
        o = rcwa1d.solve(stacks[s].flat(), lam, th, M, pol)
        err = max(np.max(np.abs(r.DEr - o["DEr"])), np.max(np.abs(r.DEt - o["DEt"])))
        assert err <= 1e-10, (i, err)
        if s % 2 == 0:
            assert abs(r.reflectance() + r.transmittance() - 1.0) <= 1e-11, i
    b.close()


def test_fft_settings_rejected_as_reference():
    from stablercwa_amd import rcwa
    from stablercwa_amd.types import FourierMethod, Layer, Stack

    def stack(samples, method=FourierMethod.FFT):
        return Stack(1.0, [Layer.grating(0.3, [0.0, 0.2, 0.5], [1.5, 1.0], method, samples)], 1.45, 0.5)
    with pytest.raises(ValueError, match=r"fft_samples must be at least 4\*M\+1"):
        rcwa.solve(stack(8), 1.0, 10.0, 2)
    with pytest.raises(ValueError, match=r"fft_samples must be 0\.\.4194304"):
        rcwa.solve(stack(4194305), 1.0, 10.0, 2)
    bad = stack(0)
    flat = bad.flat()
    flat[-2] = 2.0                                  # an unknown Fourier method word
    from stablercwa_amd import _lib
    import ctypes
    off = np.zeros(1, np.int64); idx = np.zeros(1, np.int32)
    wl = np.array([1.0]); th = np.array([10.0]); M = np.array([2], np.int32); pol = np.zeros(1, np.int32)
    h = ctypes.c_void_p()
    code = _lib.load().srcwa_plan1d_create(_lib.default_context().handle, 1, _lib.ptr(flat), _lib.ptr(off), 1,
                                           _lib.ptr(idx), _lib.ptr(wl), _lib.ptr(th), _lib.ptr(M), _lib.ptr(pol), 0,
                                           ctypes.byref(h))
    assert code == _lib.EINVAL and b"invalid Fourier method" in _lib.load().srcwa_last_error()
