"""The mixed-precision star products (SRCWA_PLAN_MIXED): complex-FP32 LU with FP32-MFMA trailing updates, FP64
iterative refinement, FP64 fallback on a stall.

* the batched solve (srcwa_zgesv_batched_mixed) against numpy and the FP64 solve: FP64-level backward error
  ||A X - B|| / (||A|| ||X|| + ||B||) <= 1e-14 on every order class of the FP32 LU (one and several
  32-column panels, a narrow last panel, right-hand-side counts on both sides of the MFMA / VALU update),
  at least one FP32 correction (the FP32 factors alone are only good to ~1e-7), no fallback on
  well-conditioned systems; a fallback on a system too ill-conditioned for FP32 (kappa ~ 1e9), with an
  FP64-accurate answer; info = 1 on an exactly singular system
* whole solves with the mixed star products against the reference's solver compiled from its headers
  (tests/golden/large_ref.npz): planar M = 150, conical M 100-120, crossed pillars M1 = M2 = 5, 7, 9, all
  within 1e-9 of the reference's order efficiencies, the bound north_star sets for this form; the
  refinement record says what each solve needed
* each mixed result is bitwise the same alone and in a mixed batch
"""
import numpy as np
import pytest

from test_gpu_large import _jones, _select

pytestmark = pytest.mark.gpu

OFF = pytest.mark.skip(
    reason="mixed-precision star products are compiled out (SRCWA_MIXED_PRECISION 0): their first GPU run, "
           "test_zgesv_mixed_orders_and_rhs, ended in a GPU memory fault whose cause is not found; a CPU emulation "
           "of the same kernel body under AddressSanitizer runs every shape of that test cleanly")

TOL = 1e-9


@pytest.fixture(scope="module")
def mctx():
    from stablercwa_amd import _lib
    return _lib.Context(0, plan_flags=_lib.PLAN_MIXED)


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def berr(A, X, B):
    return np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X) + np.linalg.norm(B))


@OFF
def test_zgesv_mixed_orders_and_rhs(ctx):
    from stablercwa_amd import rcwa
    rng = np.random.default_rng(31)
    cases = [(1, 1), (7, 3), (31, 1), (32, 32), (33, 8), (64, 7), (100, 100), (121, 121), (150, 2), (242, 242),
             (300, 9), (482, 2), (722, 16)]
    As = [crand(rng, n, n) + n ** 0.5 * np.eye(n) for n, _ in cases]
    Bs = [crand(rng, n, r) for n, r in cases]
    mixed = rcwa.zgesv_batched(As, Bs, ctx, mixed=True)
    plain = rcwa.zgesv_batched(As, Bs, ctx)
    for (n, r), A, B, (X, info, (fb, iters, res)), (X64, info64) in zip(cases, As, Bs, mixed, plain):
        assert info == 0 and info64 == 0, n
        assert fb == 0, (n, r, "fell back to FP64")
        assert 1 <= iters <= 6, (n, r, iters)
        assert berr(A, X, B) <= 1e-14, (n, r, berr(A, X, B))
        # the refined answer is the FP64 answer to the conditioning of the system
        rel = np.max(np.abs(X - X64)) / np.max(np.abs(X64))
        assert rel <= 1e-12, (n, r, rel)
        assert 0.0 <= res <= 2.2e-16 * np.sqrt(n) * 4, (n, r, res)


@OFF
def test_zgesv_mixed_falls_back_when_fp32_cannot_converge(ctx):
    """kappa(A) ~ 1e9: FP32 corrections contract by ~kappa * 2^-24 > 1 per step, so the refinement stalls and
    the FP64 LU answers; the answer is FP64-accurate either way."""
    from stablercwa_amd import rcwa
    rng = np.random.default_rng(32)
    n = 96
    U, _ = np.linalg.qr(crand(rng, n, n))
    V, _ = np.linalg.qr(crand(rng, n, n))
    s = np.logspace(0, -9, n)
    A = (U * s) @ V.conj().T
    B = crand(rng, n, 4)
    (X, info, (fb, iters, res)), = rcwa.zgesv_batched([A], [B], ctx, mixed=True)
    assert info == 0 and fb == 1
    assert berr(A, X, B) <= 1e-14


@OFF
def test_zgesv_mixed_singular(ctx):
    from stablercwa_amd import rcwa
    rng = np.random.default_rng(33)
    A = crand(rng, 40, 40)
    A[:, 7] = 0.0
    (X, info, (fb, iters, res)), = rcwa.zgesv_batched([A], [crand(rng, 40, 2)], ctx, mixed=True)
    assert info == 1 and fb == 1


def test_mixed_flag_refused_while_compiled_out():
    """With the mixed form compiled out, asking for it is an error, never a silent FP64 run."""
    from stablercwa_amd import _lib, rcwa
    with pytest.raises(NotImplementedError):
        _lib.Context(0, plan_flags=_lib.PLAN_MIXED)
    with pytest.raises(NotImplementedError):
        rcwa.zgesv_batched([np.eye(3, dtype=complex)], [np.ones((3, 1), dtype=complex)], mixed=True)


def _check_refine(b, n):
    ref = b.refinement()
    assert ref.shape == (n, 3)
    return ref


@OFF
def test_planar_mixed_M150_against_reference(mctx):
    from stablercwa_amd import rcwa
    from stablercwa_amd.types import Stack
    cases = _select("planar", False)
    b = rcwa.PlanarBatch([Stack.from_flat(c["stack"]) for c in cases],
                         [(i, c["wavelength"], c["theta"], c["M"], c["pol"]) for i, c in enumerate(cases)], mctx)
    b.execute()
    res = b.results()
    ref = _check_refine(b, len(cases))
    for i, c in enumerate(cases):
        r = b.result(i, res)
        assert res["status"][i] == 0 and r.eig_info == 0, c["name"]
        err = max(np.max(np.abs(r.DEr - c["DEr"])), np.max(np.abs(r.DEt - c["DEt"])))
        assert err <= TOL, (c["name"], err, ref[i])
    assert np.all(ref[:, 1] > 0), "every planar solve with a grating runs FP32 corrections"
    b.close()


@OFF
def test_conical_mixed_large_M_against_reference(mctx):
    from stablercwa_amd import rcwa
    from stablercwa_amd.types import Stack
    cases = _select("conical", False)
    b = rcwa.ConicalBatch([Stack.from_flat(c["stack"]) for c in cases],
                          [(i, c["wavelength"], c["theta"], c["phi"], c["M"]) for i, c in enumerate(cases)], mctx)
    b.execute()
    res = b.results()
    ref = _check_refine(b, len(cases))
    for i, c in enumerate(cases):
        r = b.result(i, res)
        assert res["status"][i] == 0 and r.eig_info == 0, c["name"]
        err = max(np.max(np.abs(r.DEr.ravel() - c["DEr"])), np.max(np.abs(r.DEt.ravel() - c["DEt"])))
        assert err <= TOL, (c["name"], err, ref[i])
    b.close()


@OFF
def test_crossed_mixed_M5_7_9_against_reference(mctx):
    from stablercwa_amd import rcwa
    from stablercwa_amd.types import Stack2D
    cases = _select("crossed", False)
    b = rcwa.CrossedBatch([Stack2D.from_flat(c["stack"]) for c in cases],
                          [(i, c["wavelength"], c["theta"], c["phi"], c["M"], c["M2"], _jones(c))
                           for i, c in enumerate(cases)], mctx)
    b.execute()
    res = b.results()
    ref = _check_refine(b, len(cases))
    for i, c in enumerate(cases):
        r = b.result(i, res)
        assert int(res["status"][i]) == 0 and r.eig_info == 0, c["name"]
        err = max(np.max(np.abs(r.DEr - c["DEr"])), np.max(np.abs(r.DEt - c["DEt"])))
        assert err <= TOL, (c["name"], err, ref[i])
    b.close()


@OFF
def test_mixed_result_alone_equals_in_mix(mctx):
    """A conical instance solved alone and next to other, larger instances: bitwise the same."""
    from stablercwa_amd import rcwa
    from stablercwa_amd.workloads import conical_batch
    probs = conical_batch(24)
    one = rcwa.ConicalBatch([probs[3].stack], [(0, probs[3].wavelength, probs[3].theta, probs[3].phi, probs[3].M)],
                            mctx)
    one.execute()
    r1 = one.results()
    one.close()
    allb = rcwa.ConicalBatch([p.stack for p in probs],
                             [(i, p.wavelength, p.theta, p.phi, p.M) for i, p in enumerate(probs)], mctx)
    allb.execute()
    ra = allb.result(3, allb.results())
    allb.close()
    assert np.array_equal(r1["DEr"].ravel(), ra.DEr.ravel())
    assert np.array_equal(r1["DEt"].ravel(), ra.DEt.ravel())
