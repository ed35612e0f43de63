"""GPU: the reference's failure contract for planar and conical solves (no exception, NaN results,
eig_info), driven through the C-ABI plan API.

* an order grazing exactly in the incidence medium keeps energy and carries none itself
  (reference tests/test_rcwa1d.cpp:1220-1228)
* an order grazing exactly in the exit medium next to the grating makes the interface singular: every
  output is NaN, no exception (tests/test_rcwa1d.cpp:1229-1232)
* an eigensolve that does not converge within max_iter sweeps gives eig_info > 0 and NaN results; the
  incidence-only geometry of a conical solve is still reported (tests/test_rcwa1d.cpp:1169-1181,
  tests/test_rcwa1d_conical.cpp:554-563, which set EigOptions::max_iter = 1 through custom_eig)
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def grating_stack(nII=1.45):
    """The reference tests' binary grating (tests/test_rcwa1d.cpp:331-340)."""
    from stablercwa_amd.types import Layer, Stack
    p, f = 0.5747, 0.30
    return Stack(1.0, [Layer.grating(0.38, [-f * p / 2, f * p / 2, p - f * p / 2], [1.45, 1.0])], nII, p)


def conical_stack_1():
    """conical_ref_stack(1) of the reference tests (tests/conical_reference.inc:870-878)."""
    from stablercwa_amd.types import Layer, Stack
    return Stack(1.0, [Layer.grating(0.38, [0.0, 0.1722, 0.574], [1.46, 1.0])], 1.46, 0.574)


def _all_nan(*arrays):
    return all(np.all(np.isnan(np.asarray(a).view(np.float64))) for a in arrays)


def test_incidence_side_grazing_keeps_energy(ctx):
    from stablercwa_amd import rcwa
    g = grating_stack()
    for pol in (0, 1):
        r = rcwa.solve(g, g.period, 0.0, 8, rcwa.Polarization(pol), ctx)
        assert abs(r.reflectance() + r.transmittance() - 1.0) <= 1e-10
        assert r.DEr[8 + 1] == 0.0 and r.DEr[8 - 1] == 0.0


def test_incidence_side_grazing_matches_reference(ctx, ref_lib):
    from oracle import ref
    from stablercwa_amd import rcwa
    g = grating_stack()
    for pol in (0, 1):
        r = rcwa.solve(g, g.period, 0.0, 8, rcwa.Polarization(pol), ctx)
        o = ref.solve1d(ref_lib, g.flat(), g.period, 0.0, 8, pol)
        assert np.max(np.abs(r.DEr - o["DEr"])) <= 1e-10 and np.max(np.abs(r.DEt - o["DEt"])) <= 1e-10


def test_exit_side_grazing_is_all_nan(ctx):
    from stablercwa_amd import rcwa
    g = grating_stack(nII=2.0)                  # orders +-2: n^2 - kx^2 = 4 - 4 = 0 in the exit medium
    b = rcwa.PlanarBatch([g], [(0, g.period, 0.0, 8, 0)], ctx)
    b.execute()
    res = b.results()
    b.close()
    assert res["status"][0] == 1
    assert _all_nan(res["R"], res["T"], res["DEr"], res["DEt"])


def _ref_eig_info_1d(ref_lib, stack, lam, theta, M, pol, max_iter):
    N = 2 * M + 1
    R = np.zeros(2 * N); T = np.zeros(2 * N); DEr = np.zeros(N); DEt = np.zeros(N)
    counts = np.zeros(4, np.int32); met = np.zeros(2)
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    f = stack.flat()
    assert ref_lib.ref_solve1d_ex(P(f), ctypes.c_double(lam), ctypes.c_double(theta), M, pol, max_iter, 0, P(R), P(T),
                                  P(DEr), P(DEt), P(counts), P(met)) == 0
    return int(counts[0]), DEr


def _ref_eig_info_conical(ref_lib, stack, lam, theta, phi, M, max_iter):
    N = 2 * M + 1
    Jr = np.zeros(8 * N); Jt = np.zeros(8 * N); DEr = np.zeros(4 * N); DEt = np.zeros(4 * N)
    counts = np.zeros(4, np.int32); met = np.zeros(2)
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    f = stack.flat()
    assert ref_lib.ref_solve_conical_ex(P(f), ctypes.c_double(lam), ctypes.c_double(theta), ctypes.c_double(phi), M,
                                        max_iter, 0, P(Jr), P(Jt), P(DEr), P(DEt), P(counts), P(met)) == 0
    return int(counts[0]), DEr


def test_planar_unconverged_eigensolve(ctx, ref_lib):
    from stablercwa_amd import rcwa
    s = grating_stack()
    for pol in (0, 1):
        b = rcwa.PlanarBatch([s], [(0, 1.053, 66.36, 6, pol)], ctx)
        b.set_max_iter(1)
        b.execute()
        res = b.results()
        info_ref, der_ref = _ref_eig_info_1d(ref_lib, s, 1.053, 66.36, 6, pol, 1)
        assert np.all(np.isnan(der_ref))
        assert res["eig_info"][0] > 0 and res["eig_info"][0] == info_ref
        assert res["status"][0] == 1 and _all_nan(res["R"], res["T"], res["DEr"], res["DEt"])
        # the same plan with the default limit converges again
        b.set_max_iter(30)
        b.execute()
        res = b.results()
        b.close()
        assert res["eig_info"][0] == 0 and res["status"][0] == 0
        assert abs(np.sum(res["DEr"]) + np.sum(res["DEt"]) - 1.0) <= 1e-12


def test_conical_unconverged_eigensolve(ctx, ref_lib):
    from stablercwa_amd import rcwa
    s = conical_stack_1()
    good = rcwa.solve_conical(s, 1.053, 66.36, 17.0, 6, ctx)
    b = rcwa.ConicalBatch([s], [(0, 1.053, 66.36, 17.0, 6)], ctx)
    b.set_max_iter(1)
    b.execute()
    res = b.results()
    r = b.result(0, res)
    b.close()
    info_ref, der_ref = _ref_eig_info_conical(ref_lib, s, 1.053, 66.36, 17.0, 6, 1)
    assert np.all(np.isnan(der_ref))
    assert r.eig_info > 0 and r.eig_info == info_ref
    assert res["status"][0] == 1 and _all_nan(r.Jr, r.Jt, r.DEr, r.DEt)
    # what depends only on the incidence conditions is still reported, bitwise as in the good solve
    for k in ("kx", "flux_r", "flux_t", "phi_m_deg"):
        assert np.array_equal(getattr(r, k), getattr(good, k)), k
    assert np.array_equal(r.kzI, good.kzI) and np.array_equal(r.kzII, good.kzII)


def test_max_iter_rejects_out_of_range(ctx):
    from stablercwa_amd import rcwa
    b = rcwa.PlanarBatch([grating_stack()], [(0, 1.053, 66.36, 6, 0)], ctx)
    with pytest.raises(ValueError, match="max_iter"):
        b.set_max_iter(0)
    b.close()
