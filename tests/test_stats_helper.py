"""conftest.not_worse_than, the distribution check the GPU rounding-error tests use: it passes a sample
drawn from the reference's own distribution and fails one that is uniformly 30 % worse or has a heavier
tail."""
import numpy as np
import pytest

from conftest import not_worse_than


def test_same_distribution_passes():
    rng = np.random.default_rng(5)
    for seed in range(20):
        r = np.random.default_rng(seed)
        ref = np.abs(r.standard_normal(4096)) * 1e-12
        gpu = np.abs(rng.standard_normal(4096)) * 1e-12
        not_worse_than(gpu, ref)


def test_worse_mean_fails():
    rng = np.random.default_rng(6)
    ref = np.abs(rng.standard_normal(4096)) * 1e-12
    gpu = 1.3 * np.abs(rng.standard_normal(4096)) * 1e-12
    with pytest.raises(AssertionError, match="mean"):
        not_worse_than(gpu, ref)


def test_heavier_tail_fails():
    rng = np.random.default_rng(7)
    ref = np.abs(rng.standard_normal(4096)) * 1e-12
    gpu = np.abs(rng.standard_normal(4096)) * 1e-12
    gpu[:150] = 4.0 * np.percentile(ref, 99)          # 3.7 % of the points far out ...
    gpu[150:] *= (ref.sum() - gpu[:150].sum()) / gpu[150:].sum()   # ... at the reference's mean
    with pytest.raises(AssertionError, match="tail"):
        not_worse_than(gpu, ref)
