"""GPU: wavelength-invariant layer operators shared by the points of a plan.

Without SRCWA_PLAN_POINT_OPS a plan computes each grating layer's operators once per execution for every
point with the same layer record (thickness aside), period or cell, truncation and, planar, polarization
(include/srcwa.h); the reference computes them in every solve (rcwa1d.hpp:432-468).  The shared form runs
the same arithmetic on the same inputs, so each result must be the per-point form's bit for bit: across
plan chunks, under graph replay, on FFT layers, for layers that differ only in thickness, and for a layer
whose Toeplitz matrix is singular.  The compiled reference checks the shared form's values at 1e-10.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("R", "T", "DEr", "DEt", "status", "eig_info")


@pytest.fixture(scope="module")
def pctx():
    from stablercwa_amd import _lib
    return _lib.Context(0, plan_flags=_lib.PLAN_POINT_OPS)


def _same_bits(a, b, what):
    for k in KEYS:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape and x.tobytes() == y.tobytes(), (what, k)


def _planar_stacks():
    from stablercwa_amd.types import FourierMethod, Layer, Stack
    p = 0.62
    g = [0.0, 0.21, 0.40, p]
    nv = [1.0, complex(1.9, 0.02), 1.46]
    fft = Layer.grating(0.15, [0.0, 0.3, p], [2.1, 1.0], method=FourierMethod.FFT, samples=4096)
    a = Stack(1.0, [Layer.grating(0.33, g, nv), Layer.uniform(0.12, 1.7), fft], 1.46, p)
    b = Stack(1.0, [Layer.grating(0.41, g, nv), Layer.uniform(0.12, 1.7), fft], 1.46, p)   # thickness only
    return [a, b, a]                                                                        # a repeated


def _planar_problems():
    pts = [(s, wl, th, 12, pol) for s in (0, 1, 2) for wl, th in ((0.9, 5.0), (1.05, 30.0), (1.3, -12.0))
           for pol in (0, 1)]
    pts += [(0, wl, 20.0, 20, pol) for wl in (0.95, 1.1) for pol in (0, 1)]
    pts += [(1, 1.2, 0.0, 7, 0)]                                                            # a single task
    return pts


def _run(batch):
    batch.execute()
    r = batch.results()
    sets = batch.operator_sets()
    return r, sets


def test_planar_shared_equals_per_point(ctx, pctx):
    from stablercwa_amd import rcwa
    stacks, probs = _planar_stacks(), _planar_problems()
    shared = rcwa.PlanarBatch(stacks, probs, ctx)
    point = rcwa.PlanarBatch(stacks, probs, pctx)
    a, na = _run(shared)
    b, nb = _run(point)
    # two grating layers x (TE, TM) at M = 12 and at M = 20; the lone M = 7 task computes its own
    assert na == 8 and nb == 0
    assert np.all(a["status"] == 0)
    _same_bits(a, b, "shared vs per-point")
    # again: an execution recomputes the sets and must not see stale ones
    a2, _ = _run(shared)
    _same_bits(a, a2, "second execute")
    shared.close()
    point.close()


def test_planar_shared_across_chunks_and_graph(ctx, pctx):
    from stablercwa_amd import rcwa
    stacks, probs = _planar_stacks(), _planar_problems()
    one = rcwa.PlanarBatch(stacks, probs, ctx)
    ref, _ = _run(one)
    need = one.device_bytes()
    one.close()
    many = rcwa.PlanarBatch(stacks, probs, ctx, max_workspace_bytes=max(1 << 20, need // 6))
    assert many.chunks() > 1
    r, sets = _run(many)
    assert sets == 8
    _same_bits(ref, r, "chunked")
    many.close()
    g = rcwa.PlanarBatch(stacks, probs, ctx)
    g.use_graph(True)
    for _ in range(3):
        r, _ = _run(g)
        _same_bits(ref, r, "graph replay")
    g.close()


def test_planar_shared_against_reference(ctx, ref_lib):
    from oracle import ref
    from stablercwa_amd import rcwa
    stacks, probs = _planar_stacks(), _planar_problems()
    b = rcwa.PlanarBatch(stacks, probs, ctx)
    res, _ = _run(b)
    assert len(probs) == 23
    for i in (0, 7, 13, 17, 18, 21, 22):
        s, wl, th, M, pol = probs[i]
        o = ref.solve1d(ref_lib, stacks[s].flat(), wl, th, M, pol)
        d = b.result(i, res)
        assert np.max(np.abs(d.DEr - o["DEr"])) <= 1e-10 and np.max(np.abs(d.DEt - o["DEt"])) <= 1e-10, i
        assert np.max(np.abs(d.R - o["R"])) <= 1e-10 and np.max(np.abs(d.T - o["T"])) <= 1e-10, i
    b.close()


def test_planar_singular_shared_operator(ctx, pctx):
    """A TM grating whose [[eps]] is exactly singular at this truncation: every point that shares it fails as
    the per-point form does (NaN results, the same status), and the other points are unaffected."""
    from stablercwa_amd import rcwa
    from stablercwa_amd.types import Layer, Stack
    # eps = 1 on half the period, -1 on the other: the zero-order coefficient vanishes and [[eps]] at M = 0
    # is the 1 x 1 zero matrix
    bad = Stack(1.0, [Layer.grating(0.2, [0.0, 0.5, 1.0], [1.0, 1j])], 1.5, 1.0)
    ok = Stack(1.0, [Layer.grating(0.2, [0.0, 0.3, 1.0], [1.5, 1.0])], 1.5, 1.0)
    probs = [(0, 1.3, 10.0, 0, 1), (0, 1.4, 0.0, 0, 1), (1, 1.3, 10.0, 0, 1), (1, 1.4, 0.0, 0, 1)]
    a = rcwa.PlanarBatch([bad, ok], probs, ctx)
    b = rcwa.PlanarBatch([bad, ok], probs, pctx)
    ra, na = _run(a)
    rb, _ = _run(b)
    assert na == 2
    _same_bits(ra, rb, "singular")
    assert np.all(ra["status"][:2] != 0) and np.all(ra["status"][2:] == 0)
    assert np.all(np.isnan(ra["DEr"][:2]))
    a.close()
    b.close()
