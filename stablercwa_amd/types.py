"""Problem and result types of the planar, conical and crossed-grating solves.

They mirror the reference's public types (``rcwa::Layer``/``Stack``/``Result`` in
include/stablercwa/rcwa1d.hpp:80-164, ``ConicalResult`` in rcwa1d_conical.hpp:105-183, ``Layer2D``/
``Stack2D``/``Result2D`` in rcwa2d_types.hpp:39-130) and know how to write themselves into the flat
double records that cross the C-ABI (``include/srcwa.h``):

  planar stack : period, nI.re, nI.im, nII.re, nII.im, L, then per layer
                 d, kind (0 uniform | 1 grating),
                   uniform: n.re, n.im
                   grating: K, edges[K+1], (n.re, n.im)[K], fourier_method (0 analytic | 1 FFT), fft_samples
  crossed stack: Lx, Ly, nI, nII.re, nII.im, L, then per layer
                 d, kind (0 uniform | 1 patterned),
                   uniform: n.re, n.im
                   patterned: Ky, y_edges[Ky+1], then per y segment Kx, x_edges[Kx+1], (n.re, n.im)[Kx]
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np


class Polarization(enum.IntEnum):
    """TE: E parallel to the grooves (principal field Ey); TM: H parallel (principal field Hy)."""
    TE = 0
    TM = 1


class FourierMethod(enum.IntEnum):
    """How a grating layer's permittivity Fourier coefficients are formed (fourier.hpp:15)."""
    Analytic = 0
    FFT = 1


class Status2D(enum.IntEnum):
    """Outcome of a crossed-grating solve (rcwa2d_types.hpp:32-37)."""
    Ok = 0
    NumericalFailure = 1
    RayleighAnomaly = 2


@dataclass
class Layer:
    """One planar layer: a uniform film or a piecewise-constant grating (rcwa1d.hpp:80-110)."""
    d: float = 0.0
    n: complex = 1.0
    edges: List[float] = field(default_factory=list)
    n_values: List[complex] = field(default_factory=list)
    homogeneous: bool = True
    fourier_method: FourierMethod = FourierMethod.Analytic
    fft_samples: int = 0

    @staticmethod
    def uniform(thickness: float, index: complex) -> "Layer":
        return Layer(d=float(thickness), n=complex(index))

    @staticmethod
    def grating(thickness: float, boundaries: Sequence[float], indices: Sequence[complex],
                method: FourierMethod = FourierMethod.Analytic, samples: int = 0) -> "Layer":
        return Layer(d=float(thickness), edges=[float(x) for x in boundaries],
                     n_values=[complex(v) for v in indices], homogeneous=False,
                     fourier_method=FourierMethod(method), fft_samples=int(samples))


@dataclass
class Stack:
    """Top to bottom: incidence medium nI | layers | exit medium nII; period is the lateral period."""
    nI: complex = 1.0
    layers: List[Layer] = field(default_factory=list)
    nII: complex = 1.0
    period: float = 1.0

    def flat(self) -> np.ndarray:
        nI, nII = complex(self.nI), complex(self.nII)
        out = [float(self.period), nI.real, nI.imag, nII.real, nII.imag, float(len(self.layers))]
        for layer in self.layers:
            out += [float(layer.d)]
            if layer.homogeneous:
                n = complex(layer.n)
                out += [0.0, n.real, n.imag]
            else:
                K = len(layer.n_values)
                out += [1.0, float(K)] + [float(x) for x in layer.edges]
                for v in layer.n_values:
                    v = complex(v)
                    out += [v.real, v.imag]
                out += [float(int(layer.fourier_method)), float(layer.fft_samples)]
        return np.asarray(out, dtype=np.float64)

    @staticmethod
    def from_flat(rec) -> "Stack":
        """Inverse of flat(): the stack a flat planar record describes."""
        p = [float(x) for x in rec]
        st = Stack(complex(p[1], p[2]), [], complex(p[3], p[4]), p[0])
        k = 6
        for _ in range(int(p[5])):
            d, kind = p[k], int(p[k + 1])
            k += 2
            if kind == 0:
                st.layers.append(Layer.uniform(d, complex(p[k], p[k + 1])))
                k += 2
            else:
                K = int(p[k])
                edges = p[k + 1:k + K + 2]
                vals = [complex(p[k + K + 2 + 2 * i], p[k + K + 3 + 2 * i]) for i in range(K)]
                k += 1 + (K + 1) + 2 * K
                st.layers.append(Layer.grating(d, edges, vals, FourierMethod(int(p[k])), int(p[k + 1])))
                k += 2
        return st


@dataclass
class Result:
    """Planar solve result (rcwa1d.hpp:145-164). Order m sits at index m + M."""
    orders: np.ndarray
    R: np.ndarray
    T: np.ndarray
    DEr: np.ndarray
    DEt: np.ndarray
    eig_info: int = 0
    n_eig: int = 0
    n_dense_solve: int = 0
    max_eig_residual: float = 0.0

    def reflectance(self) -> float:
        return float(np.sum(self.DEr))

    def transmittance(self) -> float:
        return float(np.sum(self.DEt))


@dataclass
class Jones:
    ss: complex
    sp: complex
    ps: complex
    pp: complex


@dataclass
class ConicalResult:
    """Conical solve result (rcwa1d_conical.hpp:105-183): per-order Jones matrices and channel efficiencies.

    ``Jr``/``Jt`` are arrays of shape (N, 4) in (ss, sp, ps, pp) order; ``DEr``/``DEt`` likewise hold
    flux * |J|^2 per channel."""
    orders: np.ndarray
    kx: np.ndarray
    ky: float
    kzI: np.ndarray
    kzII: np.ndarray
    flux_r: np.ndarray
    flux_t: np.ndarray
    Jr: np.ndarray
    Jt: np.ndarray
    DEr: np.ndarray
    DEt: np.ndarray
    eig_info: int = 0
    n_schur: int = 0
    phi_m_deg: np.ndarray = None

    def flux_unitarity(self) -> np.ndarray:
        """Sum over orders of flux * J^H J for both sides, as [[uss, usp], [ups, upp]] (identity when lossless)."""
        u = np.zeros((2, 2), dtype=np.complex128)
        for J, f in ((self.Jr, self.flux_r), (self.Jt, self.flux_t)):
            for j in range(J.shape[0]):
                m = np.array([[J[j, 0], J[j, 1]], [J[j, 2], J[j, 3]]])
                u += f[j] * (m.conj().T @ m)
        return u


@dataclass
class XProfile2D:
    edges: List[float]
    n_values: List[complex]


@dataclass
class Layer2D:
    """One crossed-grating layer: uniform, or y segments each with an x profile (rcwa2d_types.hpp:45-70)."""
    d: float = 0.0
    homogeneous: bool = True
    n: complex = 1.0
    y_edges: List[float] = field(default_factory=list)
    x_profiles: List[XProfile2D] = field(default_factory=list)

    @staticmethod
    def uniform(thickness: float, index: complex) -> "Layer2D":
        return Layer2D(d=float(thickness), n=complex(index))

    @staticmethod
    def patterned(thickness: float, y_boundaries: Sequence[float], rows: Sequence[XProfile2D]) -> "Layer2D":
        return Layer2D(d=float(thickness), homogeneous=False, y_edges=[float(y) for y in y_boundaries],
                       x_profiles=list(rows))


@dataclass
class Stack2D:
    nI: float = 1.0
    nII: complex = 1.0
    layers: List[Layer2D] = field(default_factory=list)
    Lx: float = 1.0
    Ly: float = 1.0

    def flat(self) -> np.ndarray:
        nII = complex(self.nII)
        out = [float(self.Lx), float(self.Ly), float(self.nI), nII.real, nII.imag, float(len(self.layers))]
        for layer in self.layers:
            out += [float(layer.d)]
            if layer.homogeneous:
                n = complex(layer.n)
                out += [0.0, n.real, n.imag]
            else:
                out += [1.0, float(len(layer.x_profiles))] + [float(y) for y in layer.y_edges]
                for prof in layer.x_profiles:
                    out += [float(len(prof.n_values))] + [float(x) for x in prof.edges]
                    for v in prof.n_values:
                        v = complex(v)
                        out += [v.real, v.imag]
        return np.asarray(out, dtype=np.float64)

    @staticmethod
    def from_flat(rec) -> "Stack2D":
        """Inverse of flat(): the crossed-grating stack a flat record describes."""
        p = [float(x) for x in rec]
        st = Stack2D(p[2], complex(p[3], p[4]), [], p[0], p[1])
        k = 6
        for _ in range(int(p[5])):
            d, kind = p[k], int(p[k + 1])
            k += 2
            if kind == 0:
                st.layers.append(Layer2D.uniform(d, complex(p[k], p[k + 1])))
                k += 2
                continue
            K = int(p[k])
            y_edges = p[k + 1:k + K + 2]
            k += K + 2
            rows = []
            for _ in range(K):
                Kx = int(p[k])
                xe = p[k + 1:k + Kx + 2]
                vals = [complex(p[k + Kx + 2 + 2 * i], p[k + Kx + 3 + 2 * i]) for i in range(Kx)]
                rows.append(XProfile2D(xe, vals))
                k += 1 + (Kx + 1) + 2 * Kx
            st.layers.append(Layer2D.patterned(d, y_edges, rows))
        return st


def jones_from_angles_2d(psi_deg: float = 90.0, delta_deg: float = 0.0):
    """[Es, Ep] = [sin(psi) exp(i delta), cos(psi)] (rcwa2d_types.hpp:103-107)."""
    psi, delta = math.radians(psi_deg), math.radians(delta_deg)
    return (math.sin(psi) * complex(math.cos(delta), math.sin(delta)), complex(math.cos(psi), 0.0))


@dataclass
class Result2D:
    """Crossed-grating result (rcwa2d_types.hpp:112-130)."""
    status: Status2D
    M1: int
    M2: int
    N: int
    DEr: np.ndarray
    DEt: np.ndarray
    Rs: np.ndarray
    Rp: np.ndarray
    Ts: np.ndarray
    Tp: np.ndarray
    Ex_r: np.ndarray = None
    Ey_r: np.ndarray = None
    Ex_t: np.ndarray = None
    Ey_t: np.ndarray = None
    Kx: np.ndarray = None
    Ky: np.ndarray = None
    kzI: np.ndarray = None
    kzII: np.ndarray = None
    eig_info: int = 0
    message: str = ""

    def reflectance(self) -> float:
        return float(np.sum(self.DEr))

    def transmittance(self) -> float:
        return float(np.sum(self.DEt))
