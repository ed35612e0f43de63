"""The request server (continuous batching over HIP streams, admission by device-memory footprint).

* a mixed stream of planar / conical / crossed requests submitted from two threads: every result is
  bitwise the same as solving that request alone with the plan API
* a pool too small for all requests at once: requests wait for memory and still finish correctly
* invalid requests fail at submit with the reference's message; oversized requests are refused
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mix(seed=11, count=24):
    from stablercwa_amd.workloads import conical_batch, mirror_stack, pillar_batch
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        k = i % 3
        P = int(rng.integers(1, 6))
        if k == 0:
            p = conical_batch(1, seed=int(rng.integers(1 << 30)), m_lo=8, m_hi=30)[0]
            st = p.stack if i % 2 else mirror_stack()
            lam = p.wavelength if i % 2 else 1.05
            reqs.append(("planar", st, lam * (1 + 0.03 * np.arange(P)), 30.0, 0.0, p.M, i % 2))
        elif k == 1:
            p = conical_batch(1, seed=int(rng.integers(1 << 30)), m_lo=6, m_hi=16)[0]
            reqs.append(("conical", p.stack, p.wavelength * (1 + 0.02 * np.arange(P)), p.theta, p.phi, p.M, 0))
        else:
            M = int(rng.integers(2, 4))
            p = pillar_batch(1, seed=int(rng.integers(1 << 30)), orders=(M,))[0]
            reqs.append(("crossed", p.stack, p.wavelength * (1 + 0.01 * np.arange(P)), p.theta, p.phi, M, 0))
    return reqs


JONES = (0.6 + 0.1j, 0.78 + 0.0j)


def _submit(srv, r):
    kind, st, lams, th, ph, M, pol = r
    if kind == "planar":
        return srv.submit_planar(st, lams, th, M, pol)
    if kind == "conical":
        return srv.submit_conical(st, lams, th, ph, M)
    return srv.submit_crossed(st, lams, th, ph, M, M, JONES)


def _alone(r, ctx):
    from stablercwa_amd import rcwa
    kind, st, lams, th, ph, M, pol = r
    if kind == "planar":
        b = rcwa.PlanarBatch([st], [(0, float(l), th, M, pol) for l in lams], ctx)
    elif kind == "conical":
        b = rcwa.ConicalBatch([st], [(0, float(l), th, ph, M) for l in lams], ctx)
    else:
        b = rcwa.CrossedBatch([st], [(0, float(l), th, ph, M, M, JONES) for l in lams], ctx)
    b.execute()
    res = b.results()
    b.close()
    if kind == "conical":
        return res["Jr"].ravel(), res["DEr"].ravel(), res["DEt"].ravel(), res["status"]
    return res["R"], res["DEr"], res["DEt"], res["status"]


def _check_same(r, got, ctx):
    R, DEr, DEt, st = _alone(r, ctx)
    assert np.array_equal(got["status"], st)
    assert np.array_equal(got["DEr"], DEr) and np.array_equal(got["DEt"], DEt)
    assert np.array_equal(got["R"], R)


def test_server_mixed_stream_equals_alone(ctx):
    from stablercwa_amd.serve import Server
    reqs = _mix()
    ids = [None] * len(reqs)
    with Server(hbm_gib=4, slots=4, max_batch_points=64) as srv:
        def worker(part):
            for i in part:
                ids[i] = _submit(srv, reqs[i])
        th = [threading.Thread(target=worker, args=(range(k, len(reqs), 2),)) for k in (0, 1)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        results = [srv.result(i) for i in ids]
        stats = srv.stats()
    assert stats["requests"] == len(reqs) and stats["batches"] >= 3
    for r, got in zip(reqs, results):
        assert np.all(got["status"] == 0)
        assert got["latency_s"] >= got["queued_s"] >= 0.0
        _check_same(r, got, ctx)


def test_server_small_pool_waits_for_memory(ctx):
    from stablercwa_amd.serve import Server
    reqs = [r for r in _mix(seed=5, count=12) if r[0] != "crossed"]
    with Server(hbm_gib=0.03, slots=3, max_batch_points=64) as srv:
        ids = [_submit(srv, r) for r in reqs]
        results = [srv.result(i) for i in ids]
    for r, got in zip(reqs, results):
        _check_same(r, got, ctx)


def test_server_rejects_invalid_and_oversized():
    from stablercwa_amd.serve import Server
    from stablercwa_amd.workloads import mirror_stack
    st = mirror_stack()
    with Server(hbm_gib=0.01, slots=1) as srv:
        with pytest.raises(ValueError, match="incidence angle"):
            srv.submit_planar(st, [1.0], 95.0, 10)
        with pytest.raises(Exception):
            srv.submit_planar(st, np.linspace(1.0, 1.1, 64), 30.0, 150)   # needs more than 10 MiB
        rid = srv.submit_planar(st, [1.0], 30.0, 5)
        assert srv.result(rid)["status"][0] == 0
