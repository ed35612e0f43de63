"""Plans replayed from a captured HIP graph give the same bits as eager launches, replay after replay,
with other eager GPU work and synchronisations in between."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


class _S:
    def __init__(self, f):
        self.f = np.asarray(f, dtype=np.float64)

    def flat(self):
        return self.f


def _planar(cases, ctx, **kw):
    from stablercwa_amd import rcwa
    return rcwa.PlanarBatch([_S(c["stack"]) for c in cases],
                            [(i, c["wavelength"], c["theta"], c["M"], c["pol"]) for i, c in enumerate(cases)], ctx, **kw)


def _same(a, b):
    for k in ("DEr", "DEt", "status", "eig_info"):
        assert np.array_equal(a[k], b[k]), k


def test_planar_graph_replays_match_eager(ctx):
    from stablercwa_amd import rcwa
    b = _planar(load_golden("planar_ref.json")["cases"], ctx)
    b.execute()
    eager = b.results()
    side = rcwa.ConicalBatch([_S(c["stack"]) for c in load_golden("conical_tables.json")["cases"][:3]],
                             [(i, c["wavelength"], c["theta"], c["phi"], c["M"])
                              for i, c in enumerate(load_golden("conical_tables.json")["cases"][:3])], ctx)
    b.use_graph(True)
    for _ in range(4):
        b.execute()
        _same(b.results(), eager)
        assert np.array_equal(b.results()["R"], eager["R"])
        side.execute()
        side.sync()
    ms = b.stage_ms()
    assert ms[3] > 0 and ms[1] == -1
    b.use_graph(False)
    b.execute()
    _same(b.results(), eager)


def test_conical_and_crossed_graph_replays(ctx):
    from stablercwa_amd import rcwa
    cc = load_golden("conical_tables.json")["cases"]
    bc = rcwa.ConicalBatch([_S(c["stack"]) for c in cc], [(i, c["wavelength"], c["theta"], c["phi"], c["M"])
                                                          for i, c in enumerate(cc)], ctx)
    xc = [c for c in load_golden("rcwa2d_tables.json")["cases"] if c["status"] == 0][:8]
    bx = rcwa.CrossedBatch([_S(c["stack"]) for c in xc],
                           [(i, c["wavelength"], c["theta"], c["phi"], c["M1"], c["M2"],
                             (complex(c["jones"][0], c["jones"][1]), complex(c["jones"][2], c["jones"][3])))
                            for i, c in enumerate(xc)], ctx)
    bc.execute()
    bx.execute()
    ec, ex = bc.results(), bx.results()
    bc.use_graph(True)
    bx.use_graph(True)
    for _ in range(3):
        bc.execute()
        bx.execute()
        rc, rx = bc.results(), bx.results()
        assert np.array_equal(rc["Jr"], ec["Jr"]) and np.array_equal(rc["DEt"], ec["DEt"])
        assert np.array_equal(rx["R"], ex["R"]) and np.array_equal(rx["DEr"], ex["DEr"])


def test_graph_needs_one_chunk(ctx):
    b = _planar(load_golden("planar_ref.json")["cases"][:4], ctx, max_workspace_bytes=1)
    assert b.chunks() == 4
    with pytest.raises(NotImplementedError):
        b.use_graph(True)
    b.execute()
    assert np.all(b.results()["status"] == 0)
