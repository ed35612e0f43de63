import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libsrcwa_amd.so")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def ctx():
    from stablercwa_amd import _lib
    return _lib.default_context()


@pytest.fixture(scope="session")
def ref_lib():
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built (the reference tree was absent at build time)")
    return ref.load()


def max_abs_diff(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.max(np.abs(a - b))) if a.size else 0.0


def not_worse_than(gpu, ref, what=""):
    """The GPU's rounding-error distribution is not worse than the reference's on the same points.

    Per-point rounding errors of two correct FP64 solvers are independent noise of one scale (on the
    benchmark sweep the ratio of the GPU's energy defect to the reference's has median 0.97 and 90th
    percentile 6), so one point against one point decides nothing.  Two one-sided checks on the whole
    set instead: the paired mean difference is at most 3 standard errors (a false alarm in about 1 of 700
    runs when the distributions are the same), and the number of points above the reference's 99th
    percentile is within 4 standard deviations of its binomial expectation (1 % of the points)."""
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    n = gpu.size
    d = gpu - ref
    se = float(np.std(d, ddof=1) / np.sqrt(n)) if n > 1 else 0.0
    assert float(np.mean(d)) <= 3.0 * se, (what, "mean", float(np.mean(gpu)), float(np.mean(ref)), se)
    p99 = float(np.percentile(ref, 99))
    above = int(np.sum(gpu > p99))
    assert above <= 0.01 * n + 4.0 * np.sqrt(0.01 * n) + 1, (what, "tail", above, n)
