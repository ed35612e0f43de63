"""Synthetic workload generators (seeded) for the benchmark, the request server and the tests.

* ``mirror_sweep``   -- a binary grating (period 0.574, d 0.38, fill 0.3, n 1.46/1.0) on a 20-layer
                        quarter-wave mirror alternating n = 2.1 / 1.46 (d = 1.053/(4n), high index
                        first), substrate 1.46, incidence 1.0, theta 66.36 deg; TE and TM.
* ``conical_batch``  -- seeded random conical instances: M uniform in [m_lo, m_hi], 1-4 grating layers
                        each with 2-5 sorted random interior edges in one period (period 1), segment
                        index uniform in [1.3, 2.5] (+ i U[0, 0.1] for half the instances), thickness
                        uniform in [0.05, 0.5], lambda in [0.8, 2.0], theta in [0, 70] deg,
                        phi in [0, 90] deg, nI = 1.0, nII = 1.46.
* ``pillar_batch``   -- seeded square lattices (period 1) of circular pillars staircased into 2M+1 y
                        slices (half-widths at slice midpoints), M1 = M2 drawn from {5, 7, 9}
                        (N = 121 / 225 / 361): radius in [0.2, 0.4], pillar index
                        in [1.5, 3.5] + i[0, 0.05] in air, 1-3 patterned layers on substrate 1.46,
                        normal or oblique incidence (theta <= 40 deg), lambda in [1.0, 2.0].
* ``served_mix``     -- a request stream for the server: 60 % planar TE/TM (M in [20, 150]; stacks from
                        the conical generator's layer recipe, or the mirror stack), 25 % conical
                        (M in [20, 80]) and 15 % crossed pillars (M1 = M2 in [3, 9]); each request is a
                        sweep of 1-64 wavelengths around its instance's lambda; Poisson arrival times.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from .types import Layer, Layer2D, Stack, Stack2D, XProfile2D


def mirror_stack() -> Stack:
    layers = [Layer.grating(0.38, [0.0, 0.3 * 0.574, 0.574], [1.46, 1.0])]
    for k in range(20):
        n = 2.1 if k % 2 == 0 else 1.46
        layers.append(Layer.uniform(1.053 / (4 * n), n))
    return Stack(1.0, layers, 1.46, 0.574)


def mirror_sweep(points: int = 4096, M: int = 60, theta: float = 66.36):
    """(stacks, problems) for the planar TE+TM sweep; problems are (stack, lambda, theta, M, pol)."""
    lams = np.linspace(0.95, 1.15, points)
    return [mirror_stack()], [(0, float(l), theta, M, pol) for pol in (0, 1) for l in lams]


@dataclass
class ConicalProblem:
    stack: Stack
    wavelength: float
    theta: float
    phi: float
    M: int


def conical_batch(count: int = 2048, seed: int = 0xC3, m_lo: int = 20, m_hi: int = 120) -> List[ConicalProblem]:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        M = int(rng.integers(m_lo, m_hi + 1))
        lossy = (i % 2) == 1
        layers = []
        for _ in range(int(rng.integers(1, 5))):
            k = int(rng.integers(2, 6))
            interior = np.sort(rng.uniform(0.0, 1.0, k))
            edges = [0.0] + [float(x) for x in interior] + [1.0]
            vals = [complex(rng.uniform(1.3, 2.5), rng.uniform(0.0, 0.1) if lossy else 0.0) for _ in range(k + 1)]
            layers.append(Layer.grating(float(rng.uniform(0.05, 0.5)), edges, vals))
        st = Stack(1.0, layers, 1.46, 1.0)
        out.append(ConicalProblem(st, float(rng.uniform(0.8, 2.0)), float(rng.uniform(0.0, 70.0)),
                                  float(rng.uniform(0.0, 90.0)), M))
    return out


def pillar_layer(d: float, radius: float, n_pillar: complex, slices: int) -> Layer2D:
    """A circular pillar of the given radius at the cell centre, staircased into `slices` y slices."""
    y_edges = [k / slices for k in range(slices + 1)]
    rows = []
    for k in range(slices):
        yc = (k + 0.5) / slices - 0.5
        half = math.sqrt(max(radius * radius - yc * yc, 0.0))
        if half <= 0.0:
            rows.append(XProfile2D([0.0, 1.0], [1.0]))
        else:
            rows.append(XProfile2D([0.0, 0.5 - half, 0.5 + half, 1.0], [1.0, n_pillar, 1.0]))
    return Layer2D.patterned(d, y_edges, rows)


@dataclass
class CrossedProblem:
    stack: Stack2D
    wavelength: float
    theta: float
    phi: float
    M1: int
    M2: int


def pillar_batch(count: int = 512, seed: int = 0xC4, orders=(5, 7, 9), oblique_fraction: float = 0.5) -> List[CrossedProblem]:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        M = int(orders[int(rng.integers(0, len(orders)))])        # M1 = M2 in {5, 7, 9}: N = 121 / 225 / 361
        slices = 2 * M + 1
        layers = []
        for _ in range(int(rng.integers(1, 4))):
            layers.append(pillar_layer(float(rng.uniform(0.1, 0.5)), float(rng.uniform(0.2, 0.4)),
                                       complex(rng.uniform(1.5, 3.5), rng.uniform(0.0, 0.05)), slices))
        st = Stack2D(1.0, 1.46, layers, 1.0, 1.0)
        theta = 0.0 if rng.random() > oblique_fraction else float(rng.uniform(0.0, 40.0))
        lam = float(rng.uniform(1.0, 2.0))
        if lam == 1.0:
            lam = 1.0 + 1e-9
        out.append(CrossedProblem(st, lam, theta, float(rng.uniform(0.0, 90.0)), M, M))
    return out


@dataclass
class ServedRequest:
    kind: str                 # "planar" | "conical" | "crossed"
    stack: object
    wavelengths: np.ndarray
    theta: float
    phi: float
    M: int
    pol: int = 0
    arrival_s: float = 0.0


def served_mix(count: int = 400, seed: int = 0xC5, rate_per_s: float = 50.0,
               max_points: int = 64) -> List[ServedRequest]:
    """Seeded mixed request stream with exponential inter-arrival times at rate_per_s."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for i in range(count):
        t += float(rng.exponential(1.0 / rate_per_s))
        u = rng.random()
        points = int(rng.integers(1, max_points + 1))
        if u < 0.60:
            M = int(rng.integers(20, 151))
            if rng.random() < 0.25:
                st, lam, theta = mirror_stack(), 1.05, 66.36
            else:
                p = conical_batch(1, seed=int(rng.integers(1 << 31)), m_lo=M, m_hi=M)[0]
                st, lam, theta = p.stack, p.wavelength, p.theta
            lams = lam * (1.0 + 0.05 * np.linspace(-1.0, 1.0, points)) if points > 1 else np.array([lam])
            out.append(ServedRequest("planar", st, lams, theta, 0.0, M, int(rng.integers(0, 2)), t))
        elif u < 0.85:
            M = int(rng.integers(20, 81))
            p = conical_batch(1, seed=int(rng.integers(1 << 31)), m_lo=M, m_hi=M)[0]
            lams = p.wavelength * (1.0 + 0.05 * np.linspace(-1.0, 1.0, points)) if points > 1 else np.array([p.wavelength])
            out.append(ServedRequest("conical", p.stack, lams, p.theta, p.phi, M, 0, t))
        else:
            M = int(rng.integers(3, 10))
            p = pillar_batch(1, seed=int(rng.integers(1 << 31)), orders=(M,))[0]
            lams = p.wavelength * (1.0 + 0.02 * np.linspace(-1.0, 1.0, points)) if points > 1 else np.array([p.wavelength])
            lams = np.where(np.abs(lams - 1.0) < 1e-9, 1.0 + 1e-6, lams)     # keep clear of the lambda = period anomaly
            out.append(ServedRequest("crossed", p.stack, lams, p.theta, p.phi, M, 0, t))
    return out
