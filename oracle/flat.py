"""Decoder of the flat double stack records (see include/srcwa.h). Test infrastructure only."""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np


@dataclass
class OLayer:
    d: float
    homogeneous: bool
    n: complex = 1.0
    edges: List[float] = field(default_factory=list)
    values: List[complex] = field(default_factory=list)
    fourier_method: int = 0
    fft_samples: int = 0


@dataclass
class OStack:
    period: float
    nI: complex
    nII: complex
    layers: List[OLayer]


def decode1d(p) -> OStack:
    p = np.asarray(p, dtype=np.float64)
    L = int(p[5])
    k = 6
    layers = []
    for _ in range(L):
        d = float(p[k]); kind = int(p[k + 1]); k += 2
        if kind == 0:
            layers.append(OLayer(d, True, complex(p[k], p[k + 1])))
            k += 2
        else:
            K = int(p[k]); k += 1
            edges = [float(x) for x in p[k:k + K + 1]]; k += K + 1
            vals = [complex(p[k + 2 * i], p[k + 2 * i + 1]) for i in range(K)]; k += 2 * K
            method = int(p[k]); samples = int(p[k + 1]); k += 2
            layers.append(OLayer(d, False, 1.0, edges, vals, method, samples))
    return OStack(float(p[0]), complex(p[1], p[2]), complex(p[3], p[4]), layers)


@dataclass
class OXProfile:
    edges: List[float]
    values: List[complex]


@dataclass
class OLayer2D:
    d: float
    homogeneous: bool
    n: complex = 1.0
    y_edges: List[float] = field(default_factory=list)
    rows: List[OXProfile] = field(default_factory=list)


@dataclass
class OStack2D:
    Lx: float
    Ly: float
    nI: float
    nII: complex
    layers: List[OLayer2D]


def decode2d(p) -> OStack2D:
    p = np.asarray(p, dtype=np.float64)
    L = int(p[5])
    k = 6
    layers = []
    for _ in range(L):
        d = float(p[k]); kind = int(p[k + 1]); k += 2
        if kind == 0:
            layers.append(OLayer2D(d, True, complex(p[k], p[k + 1])))
            k += 2
        else:
            Ky = int(p[k]); k += 1
            ye = [float(y) for y in p[k:k + Ky + 1]]; k += Ky + 1
            rows = []
            for _ in range(Ky):
                Kx = int(p[k]); k += 1
                xe = [float(x) for x in p[k:k + Kx + 1]]; k += Kx + 1
                vals = [complex(p[k + 2 * i], p[k + 2 * i + 1]) for i in range(Kx)]; k += 2 * Kx
                rows.append(OXProfile(xe, vals))
            layers.append(OLayer2D(d, False, 1.0, ye, rows))
    return OStack2D(float(p[0]), float(p[1]), float(p[2]), complex(p[3], p[4]), layers)
