"""Python mirror of the reference's solver entry points, backed by libsrcwa_amd.so on the GPU.

    from stablercwa_amd import rcwa
    r = rcwa.solve(stack, 1.053, 66.36, 20, rcwa.Polarization.TE)      # rcwa1d.hpp:1077
    rs = rcwa.sweep(stack, [rcwa.Point(l, 66.36) for l in lams], 20)  # rcwa1d.hpp:1084

``solve``/``sweep`` keep the reference's names, argument meaning and error behaviour: invalid input
raises ``ValueError`` (the reference's std::invalid_argument), a numerical failure returns NaN arrays
with ``eig_info`` set.  ``PlanarBatch`` is the batched, device-resident form used by the benchmark and
the request server: inputs are uploaded once, ``execute`` re-runs every solve on the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence

import ctypes
import numpy as np

from . import _lib
from .types import (ConicalResult, FourierMethod, Jones, Layer, Layer2D, Polarization, Result, Result2D,
                    Stack, Stack2D, Status2D, XProfile2D, jones_from_angles_2d)

__all__ = ["solve", "sweep", "Point", "PlanarBatch", "Layer", "Stack", "Polarization", "FourierMethod",
           "Result", "solve_conical", "sweep_conical", "ConicalPoint", "ConicalBatch", "ConicalResult",
           "solve2d", "CrossedBatch", "Stack2D", "Layer2D", "XProfile2D", "Result2D", "Status2D",
           "jones_from_angles_2d", "zgeev_batched", "zgesv_batched", "zgemm_batched"]

ST_OK, ST_NUMERICAL, ST_WORKSPACE, ST_RAYLEIGH, ST_SCHUR = 0, 1, 2, 3, 4


@dataclass
class Point:
    wavelength: float = 1.0
    theta_deg: float = 0.0


def _pack_stacks(stacks: Sequence[Stack]):
    recs = [s.flat() for s in stacks]
    offsets = np.zeros(len(recs), dtype=np.int64)
    pos = 0
    for i, r in enumerate(recs):
        offsets[i] = pos
        pos += r.size
    return np.concatenate(recs).astype(np.float64), offsets


class PlanarBatch:
    """A device-resident batch of planar solves (one launch plan).

    problems: sequence of (stack_index, wavelength, theta_deg, M, pol)."""

    def __init__(self, stacks: Sequence[Stack], problems, ctx: _lib.Context | None = None,
                 max_workspace_bytes: int = 0):
        self.ctx = ctx or _lib.default_context()
        self.lib = _lib.load()
        self._flat, self._offsets = _pack_stacks(stacks)
        P = len(problems)
        self.stack_index = np.array([p[0] for p in problems], dtype=np.int32)
        self.wavelength = np.array([p[1] for p in problems], dtype=np.float64)
        self.theta = np.array([p[2] for p in problems], dtype=np.float64)
        self.M = np.array([p[3] for p in problems], dtype=np.int32)
        self.pol = np.array([int(p[4]) for p in problems], dtype=np.int32)
        self.N = 2 * self.M + 1
        self.order_offsets = np.concatenate([[0], np.cumsum(self.N)]).astype(np.int64)
        h = ctypes.c_void_p()
        _lib.check(self.lib.srcwa_plan1d_create(
            self.ctx.handle, len(stacks), _lib.ptr(self._flat), _lib.ptr(self._offsets), P,
            _lib.ptr(self.stack_index), _lib.ptr(self.wavelength), _lib.ptr(self.theta), _lib.ptr(self.M),
            _lib.ptr(self.pol), int(max_workspace_bytes), ctypes.byref(h)))
        self.handle = h
        self.n = P

    def execute(self) -> None:
        _lib.check(self.lib.srcwa_plan_execute(self.handle))

    def use_graph(self, enable: bool = True) -> None:
        """Capture the launch sequence into a HIP graph that execute() replays (one-chunk plans)."""
        _lib.check(self.lib.srcwa_plan_set_graph(self.handle, 1 if enable else 0))

    def set_max_iter(self, max_iter: int) -> None:
        """QR sweeps per eigenvalue of every layer eigensolve (EigOptions::max_iter; default 30)."""
        _lib.check(self.lib.srcwa_plan_set_max_iter(self.handle, int(max_iter)))

    def sync(self) -> None:
        _lib.check(self.lib.srcwa_plan_sync(self.handle))

    def results(self):
        o = int(self.order_offsets[-1])
        R = np.empty(2 * o); T = np.empty(2 * o); DEr = np.empty(o); DEt = np.empty(o)
        st = np.empty(self.n, dtype=np.int32); info = np.empty(self.n, dtype=np.int32)
        _lib.check(self.lib.srcwa_plan_results(self.handle, _lib.ptr(R), _lib.ptr(T), _lib.ptr(DEr), _lib.ptr(DEt),
                                               _lib.ptr(st), _lib.ptr(info)))
        return dict(R=R.view(np.complex128), T=T.view(np.complex128), DEr=DEr, DEt=DEt, status=st, eig_info=info)

    def result(self, i: int, res=None) -> Result:
        res = res or self.results()
        a, b = int(self.order_offsets[i]), int(self.order_offsets[i + 1])
        M = int(self.M[i])
        return Result(orders=np.arange(-M, M + 1), R=res["R"][a:b].copy(), T=res["T"][a:b].copy(),
                      DEr=res["DEr"][a:b].copy(), DEt=res["DEt"][a:b].copy(), eig_info=int(res["eig_info"][i]))

    def diagnostics(self) -> np.ndarray:
        """Per problem [max_eig_residual, max_solve_backward_error, n_eig, n_dense_solve, n_schur] of the last
        execute (residuals need a context created with plan_flags=PLAN_DIAGNOSTICS)."""
        out = np.zeros((self.n, 5))
        _lib.check(self.lib.srcwa_plan_diagnostics(self.handle, _lib.ptr(out)))
        return out

    def refinement(self) -> np.ndarray:
        """Per problem [FP64 fallbacks, FP32 correction steps, largest final refined residual] of the last
        execute's mixed-precision solves (a context created with plan_flags=PLAN_MIXED; zeros otherwise)."""
        out = np.zeros((self.n, 3))
        _lib.check(self.lib.srcwa_plan_refinement(self.handle, _lib.ptr(out)))
        return out

    # ---- near fields (a context created with plan_flags=PLAN_FIELDS; rcwa1d_fields.hpp:55-206) ----
    def field_available(self, i: int) -> bool:
        """PlanarFieldContext::available() of problem i after execute."""
        v = ctypes.c_int(0)
        _lib.check(self.lib.srcwa_plan1d_field_available(self.handle, int(i), ctypes.byref(v)))
        return bool(v.value)

    def field_at(self, i: int, zs) -> np.ndarray:
        """Harmonic fields of problem i at the depths zs: array [6, N, nz] complex (Ex, Ey, Ez, Hx, Hy, Hz; H is
        Z0 H), as HarmonicFields::components (rcwa_fields.hpp:108-128)."""
        z = np.ascontiguousarray(np.atleast_1d(np.asarray(zs, dtype=np.float64)))
        N = int(self.N[i])
        out = np.zeros(12 * N * z.size)
        _lib.check(self.lib.srcwa_plan1d_field_at(self.handle, int(i), int(z.size), _lib.ptr(z), _lib.ptr(out)))
        return out.view(np.complex128).reshape(6, z.size, N).transpose(0, 2, 1)

    def field_map(self, i: int, zs, nx: int = 256, y: float = 0.0) -> np.ndarray:
        """The fields of problem i on the grid x_k = k period / nx (k < nx) at the depths zs: array [6, nz, nx]
        complex, as FieldGridXZ::components (data()[iz nx + ix], rcwa_fields.hpp:130-157)."""
        z = np.ascontiguousarray(np.atleast_1d(np.asarray(zs, dtype=np.float64)))
        out = np.zeros(12 * max(int(nx), 0) * z.size)
        _lib.check(self.lib.srcwa_plan1d_field_map(self.handle, int(i), int(z.size), _lib.ptr(z), int(nx), float(y),
                                                   _lib.ptr(out)))
        return out.view(np.complex128).reshape(6, z.size, max(int(nx), 0))

    def stage_ms(self) -> np.ndarray:
        ms = np.zeros(4)
        _lib.check(self.lib.srcwa_plan_stage_ms(self.handle, _lib.ptr(ms)))
        return ms

    def flops(self) -> np.ndarray:
        f = np.zeros(2)
        _lib.check(self.lib.srcwa_plan_flops(self.handle, _lib.ptr(f)))
        return f

    def device_bytes(self) -> int:
        return int(self.lib.srcwa_plan_device_bytes(self.handle))

    def operator_sets(self) -> int:
        """Shared operator sets of the plan: layers whose wavelength-invariant operators several points share,
        computed once per execute (0 with PLAN_POINT_OPS or diagnostics)."""
        return int(self.lib.srcwa_plan_operator_sets(self.handle))

    def chunks(self) -> int:
        return int(self.lib.srcwa_plan_chunks(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.srcwa_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve(stack: Stack, wavelength: float, theta_deg: float, M: int, pol: Polarization = Polarization.TE,
          ctx: _lib.Context | None = None) -> Result:
    """One planar solve (rcwa::solve, rcwa1d.hpp:1077)."""
    b = PlanarBatch([stack], [(0, wavelength, theta_deg, M, pol)], ctx)
    b.execute()
    r = b.result(0)
    b.close()
    return r


class PlanarFieldContext:
    """The near-field capture of one planar solve (PlanarFieldContext, rcwa1d_fields.hpp:16-47), kept on the
    device by its plan.  field_at / field_map raise FieldUnavailable when the solve failed."""

    def __init__(self, batch: PlanarBatch, available: bool):
        self._batch = batch
        self._available = available

    def available(self) -> bool:
        return self._available

    def field_at(self, zs) -> np.ndarray:
        """[6, N, nz] complex harmonic fields (Ex, Ey, Ez, Hx, Hy, Hz; H is Z0 H) at the depths zs."""
        return self._batch.field_at(0, zs)

    def field_map(self, zs, nx: int = 256, y: float = 0.0) -> np.ndarray:
        """[6, nz, nx] complex fields on the grid x_k = k period / nx at the depths zs."""
        return self._batch.field_map(0, zs, nx, y)


_field_contexts = {}


def solve_with_fields(stack: Stack, wavelength: float, theta_deg: float, M: int,
                      pol: Polarization = Polarization.TE, ctx: _lib.Context | None = None):
    """One planar solve and its near-field capture (rcwa::solve_with_fields, rcwa1d_fields.hpp:55-104): returns
    (Result, PlanarFieldContext).  ctx, when given, must have been created with plan_flags including PLAN_FIELDS."""
    if ctx is None:
        ctx = _field_contexts.get(0)
        if ctx is None:
            ctx = _field_contexts[0] = _lib.Context(0, plan_flags=_lib.PLAN_FIELDS)
    b = PlanarBatch([stack], [(0, wavelength, theta_deg, M, pol)], ctx)
    b.execute()
    r = b.result(0)
    return r, PlanarFieldContext(b, b.field_available(0))


def sweep(stack: Stack, points: Sequence[Point], M: int, pol: Polarization = Polarization.TE,
          ctx: _lib.Context | None = None) -> List[Result]:
    """Independent solves over points, one batched launch (rcwa::sweep, rcwa1d.hpp:1084-1091)."""
    b = PlanarBatch([stack], [(0, p.wavelength, p.theta_deg, M, pol) for p in points], ctx)
    b.execute()
    res = b.results()
    out = [b.result(i, res) for i in range(len(points))]
    b.close()
    return out


@dataclass
class ConicalPoint:
    wavelength: float = 1.0
    theta_deg: float = 0.0
    phi_deg: float = 0.0


class ConicalBatch(PlanarBatch):
    """A device-resident batch of conical solves; problems are (stack_index, wavelength, theta, phi, M)."""

    def __init__(self, stacks: Sequence[Stack], problems, ctx: _lib.Context | None = None,
                 max_workspace_bytes: int = 0):
        self.ctx = ctx or _lib.default_context()
        self.lib = _lib.load()
        self._flat, self._offsets = _pack_stacks(stacks)
        P = len(problems)
        self.stack_index = np.array([p[0] for p in problems], dtype=np.int32)
        self.wavelength = np.array([p[1] for p in problems], dtype=np.float64)
        self.theta = np.array([p[2] for p in problems], dtype=np.float64)
        self.phi = np.array([p[3] for p in problems], dtype=np.float64)
        self.M = np.array([p[4] for p in problems], dtype=np.int32)
        self.N = 2 * self.M + 1
        self.order_offsets = np.concatenate([[0], np.cumsum(self.N)]).astype(np.int64)
        h = ctypes.c_void_p()
        _lib.check(self.lib.srcwa_plan_conical_create(
            self.ctx.handle, len(stacks), _lib.ptr(self._flat), _lib.ptr(self._offsets), P,
            _lib.ptr(self.stack_index), _lib.ptr(self.wavelength), _lib.ptr(self.theta), _lib.ptr(self.phi),
            _lib.ptr(self.M), int(max_workspace_bytes), ctypes.byref(h)))
        self.handle = h
        self.n = P

    def results(self):
        o = 4 * int(self.order_offsets[-1])
        R = np.empty(2 * o); T = np.empty(2 * o); DEr = np.empty(o); DEt = np.empty(o)
        st = np.empty(self.n, dtype=np.int32); info = np.empty(self.n, dtype=np.int32)
        _lib.check(self.lib.srcwa_plan_results(self.handle, _lib.ptr(R), _lib.ptr(T), _lib.ptr(DEr), _lib.ptr(DEt),
                                               _lib.ptr(st), _lib.ptr(info)))
        geo = np.empty(8 * (o // 4))
        _lib.check(self.lib.srcwa_plan_geometry(self.handle, _lib.ptr(geo)))
        return dict(Jr=R.view(np.complex128).reshape(-1, 4), Jt=T.view(np.complex128).reshape(-1, 4),
                    DEr=DEr.reshape(-1, 4), DEt=DEt.reshape(-1, 4), status=st, eig_info=info,
                    geometry=geo.reshape(-1, 8))

    def result(self, i: int, res=None) -> ConicalResult:
        res = res or self.results()
        a, b = int(self.order_offsets[i]), int(self.order_offsets[i + 1])
        M = int(self.M[i])
        g = res["geometry"][a:b]
        th, ph = np.radians(self.theta[i]), np.radians(self.phi[i])
        nI = float(self._flat[int(self._offsets[self.stack_index[i]]) + 1])
        return ConicalResult(orders=np.arange(-M, M + 1), kx=g[:, 0].copy(), ky=nI * np.sin(th) * np.sin(ph),
                             kzI=g[:, 1] + 1j * g[:, 2], kzII=g[:, 3] + 1j * g[:, 4], flux_r=g[:, 5].copy(),
                             flux_t=g[:, 6].copy(), Jr=res["Jr"][a:b].copy(), Jt=res["Jt"][a:b].copy(),
                             DEr=res["DEr"][a:b].copy(), DEt=res["DEt"][a:b].copy(),
                             eig_info=int(res["eig_info"][i]), phi_m_deg=g[:, 7].copy())

    # ---- near fields (a context created with plan_flags=PLAN_FIELDS; rcwa1d_conical_fields.hpp:56-262) ----
    def field_at(self, i: int, zs, jones=(1.0, 0.0)) -> np.ndarray:
        """Harmonic fields of problem i at the depths zs for the incident Jones vector (Es, Ep), normalized first:
        array [6, N, nz] complex (Ex, Ey, Ez, Hx, Hy, Hz; H is Z0 H)."""
        z = np.ascontiguousarray(np.atleast_1d(np.asarray(zs, dtype=np.float64)))
        j = np.array([complex(jones[0]).real, complex(jones[0]).imag, complex(jones[1]).real, complex(jones[1]).imag])
        N = int(self.N[i])
        out = np.zeros(12 * N * z.size)
        _lib.check(self.lib.srcwa_plan_conical_field_at(self.handle, int(i), int(z.size), _lib.ptr(z), _lib.ptr(j),
                                                        _lib.ptr(out)))
        return out.view(np.complex128).reshape(6, z.size, N).transpose(0, 2, 1)

    def field_map(self, i: int, zs, nx: int = 256, y: float = 0.0, jones=(1.0, 0.0)) -> np.ndarray:
        """The fields of problem i on the grid x_k = k period / nx at height y: array [6, nz, nx] complex."""
        z = np.ascontiguousarray(np.atleast_1d(np.asarray(zs, dtype=np.float64)))
        j = np.array([complex(jones[0]).real, complex(jones[0]).imag, complex(jones[1]).real, complex(jones[1]).imag])
        out = np.zeros(12 * max(int(nx), 0) * z.size)
        _lib.check(self.lib.srcwa_plan_conical_field_map(self.handle, int(i), int(z.size), _lib.ptr(z), int(nx),
                                                         float(y), _lib.ptr(j), _lib.ptr(out)))
        return out.view(np.complex128).reshape(6, z.size, max(int(nx), 0))


def solve_conical(stack: Stack, wavelength: float, theta_deg: float, phi_deg: float, M: int,
                  ctx: _lib.Context | None = None) -> ConicalResult:
    """One conical solve (rcwa::solve_conical, rcwa1d_conical.hpp:985)."""
    b = ConicalBatch([stack], [(0, wavelength, theta_deg, phi_deg, M)], ctx)
    b.execute()
    r = b.result(0)
    b.close()
    return r


class ConicalFieldContext:
    """The near-field capture of one conical solve for the s and p incidence (ConicalFieldContext,
    rcwa1d_conical_fields.hpp:16-52), kept on the device by its plan; a Jones vector is chosen per call."""

    def __init__(self, batch: ConicalBatch, available: bool):
        self._batch = batch
        self._available = available

    def available(self) -> bool:
        return self._available

    def field_at(self, zs, jones=(1.0, 0.0)) -> np.ndarray:
        return self._batch.field_at(0, zs, jones)

    def field_map(self, zs, nx: int = 256, y: float = 0.0, jones=(1.0, 0.0)) -> np.ndarray:
        return self._batch.field_map(0, zs, nx, y, jones)


def solve_conical_with_fields(stack: Stack, wavelength: float, theta_deg: float, phi_deg: float, M: int,
                              ctx: _lib.Context | None = None):
    """One conical solve and its near-field capture (rcwa::solve_conical_with_fields,
    rcwa1d_conical_fields.hpp:56-98): returns (ConicalResult, ConicalFieldContext)."""
    if ctx is None:
        ctx = _field_contexts.get(0)
        if ctx is None:
            ctx = _field_contexts[0] = _lib.Context(0, plan_flags=_lib.PLAN_FIELDS)
    b = ConicalBatch([stack], [(0, wavelength, theta_deg, phi_deg, M)], ctx)
    b.execute()
    r = b.result(0)
    return r, ConicalFieldContext(b, b.field_available(0))


def sweep_conical(stack: Stack, points: Sequence[ConicalPoint], M: int,
                  ctx: _lib.Context | None = None) -> List[ConicalResult]:
    """Independent conical solves over points in one batched launch (rcwa1d_conical.hpp:992-999)."""
    b = ConicalBatch([stack], [(0, p.wavelength, p.theta_deg, p.phi_deg, M) for p in points], ctx)
    b.execute()
    res = b.results()
    out = [b.result(i, res) for i in range(len(points))]
    b.close()
    return out


# ---- crossed gratings ------------------------------------------------------------------------------
_STATUS2D = {ST_OK: Status2D.Ok, ST_NUMERICAL: Status2D.NumericalFailure, ST_RAYLEIGH: Status2D.RayleighAnomaly}

class CrossedBatch(PlanarBatch):
    """A device-resident batch of crossed-grating solves.

    problems: sequence of (stack_index, wavelength, theta_deg, phi_deg, M1, M2, jones[, factorization]) with
    jones = (Es, Ep) complex and factorization "li" (default) or "laurent" (rcwa2d_types.hpp:27-29)."""

    def __init__(self, stacks: Sequence[Stack2D], problems, ctx: _lib.Context | None = None,
                 max_workspace_bytes: int = 0):
        self.ctx = ctx or _lib.default_context()
        self.lib = _lib.load()
        self._flat, self._offsets = _pack_stacks(stacks)
        P = len(problems)
        self.stack_index = np.array([p[0] for p in problems], dtype=np.int32)
        self.wavelength = np.array([p[1] for p in problems], dtype=np.float64)
        self.theta = np.array([p[2] for p in problems], dtype=np.float64)
        self.phi = np.array([p[3] for p in problems], dtype=np.float64)
        self.M1 = np.array([p[4] for p in problems], dtype=np.int32)
        self.M2 = np.array([p[5] for p in problems], dtype=np.int32)
        jones = []
        for p in problems:
            es, ep = complex(p[6][0]), complex(p[6][1])
            jones += [es.real, es.imag, ep.real, ep.imag]
        self.jones = np.array(jones, dtype=np.float64)
        fz = []
        for p in problems:
            f = p[7] if len(p) > 7 else "li"
            if f not in ("li", "laurent"):
                raise ValueError("rcwa2d: factorization must be 'li' or 'laurent'")
            fz.append(0 if f == "li" else 1)
        self.factorization = np.array(fz, dtype=np.int32)
        self.N = (2 * self.M1.astype(np.int64) + 1) * (2 * self.M2.astype(np.int64) + 1)
        self.order_offsets = np.concatenate([[0], np.cumsum(self.N)]).astype(np.int64)
        h = ctypes.c_void_p()
        _lib.check(self.lib.srcwa_plan2d_create(
            self.ctx.handle, len(stacks), _lib.ptr(self._flat), _lib.ptr(self._offsets), P,
            _lib.ptr(self.stack_index), _lib.ptr(self.wavelength), _lib.ptr(self.theta), _lib.ptr(self.phi),
            _lib.ptr(self.M1), _lib.ptr(self.M2), _lib.ptr(self.jones), _lib.ptr(self.factorization),
            int(max_workspace_bytes), ctypes.byref(h)))
        self.handle = h
        self.n = P

    def results(self):
        o = int(self.order_offsets[-1])
        R = np.empty(8 * o); T = np.empty(8 * o); DEr = np.empty(o); DEt = np.empty(o)
        st = np.empty(self.n, dtype=np.int32); info = np.empty(self.n, dtype=np.int32)
        _lib.check(self.lib.srcwa_plan_results(self.handle, _lib.ptr(R), _lib.ptr(T), _lib.ptr(DEr), _lib.ptr(DEt),
                                               _lib.ptr(st), _lib.ptr(info)))
        geo = np.empty(8 * o)
        _lib.check(self.lib.srcwa_plan_geometry(self.handle, _lib.ptr(geo)))
        return dict(R=R.view(np.complex128), T=T.view(np.complex128), DEr=DEr, DEt=DEt, status=st, eig_info=info,
                    geometry=geo.reshape(-1, 8))

    def result(self, i: int, res=None) -> Result2D:
        res = res or self.results()
        a, b = int(self.order_offsets[i]), int(self.order_offsets[i + 1])
        N = b - a
        R, T = res["R"][4 * a:4 * b], res["T"][4 * a:4 * b]
        g = res["geometry"][a:b]
        code = int(res["status"][i])
        if code not in _STATUS2D:
            raise _lib.SrcwaError(code, "rcwa2d: device workspace exhausted")
        return Result2D(status=_STATUS2D[code], M1=int(self.M1[i]), M2=int(self.M2[i]), N=N,
                        DEr=res["DEr"][a:b].copy(), DEt=res["DEt"][a:b].copy(),
                        Rs=R[:N].copy(), Rp=R[N:2 * N].copy(), Ts=T[:N].copy(), Tp=T[N:2 * N].copy(),
                        Ex_r=R[2 * N:3 * N].copy(), Ey_r=R[3 * N:].copy(),
                        Ex_t=T[2 * N:3 * N].copy(), Ey_t=T[3 * N:].copy(), eig_info=int(res["eig_info"][i]),
                        Kx=g[:, 0].copy(), Ky=g[:, 1].copy(), kzI=g[:, 2] + 1j * g[:, 3], kzII=g[:, 4] + 1j * g[:, 5])


def solve2d(stack: Stack2D, wavelength: float, theta_deg: float, phi_deg: float, M1: int, M2: int,
            jones=None, factorization: str = "li", ctx: _lib.Context | None = None) -> Result2D:
    """One crossed-grating solve (rcwa::solve2d, rcwa2d.hpp:1355-1411).  jones defaults to s polarisation."""
    jones = jones if jones is not None else jones_from_angles_2d(90.0)
    b = CrossedBatch([stack], [(0, wavelength, theta_deg, phi_deg, M1, M2, jones, factorization)], ctx)
    b.execute()
    r = b.result(0)
    b.close()
    return r


# ---- hook-shaped dense batches ---------------------------------------------------------------------
def _pack(mats):
    return np.concatenate([np.asfortranarray(m, dtype=np.complex128).ravel(order="F") for m in mats]).view(np.float64)


def zgeev_batched(mats, ctx: _lib.Context | None = None):
    """Eigenpairs of each square matrix (custom_eig shape): returns [(w, V, info)]."""
    ctx = ctx or _lib.default_context()
    n = np.array([m.shape[0] for m in mats], dtype=np.int32)
    A = _pack(mats)
    tot = int(np.sum(n.astype(np.int64) ** 2))
    V = np.zeros(2 * tot); w = np.zeros(2 * int(n.sum())); info = np.zeros(len(mats), dtype=np.int32)
    _lib.check(_lib.load().srcwa_zgeev_batched(ctx.handle, len(mats), _lib.ptr(n), _lib.ptr(A), _lib.ptr(w),
                                               _lib.ptr(V), _lib.ptr(info)))
    Vc, wc = V.view(np.complex128), w.view(np.complex128)
    out, a, b = [], 0, 0
    for i, k in enumerate(n):
        out.append((wc[b:b + k].copy(), Vc[a:a + k * k].reshape((k, k), order="F").copy(), int(info[i])))
        a += k * k; b += k
    return out


def zgesv_batched(As, Bs, ctx: _lib.Context | None = None, mixed: bool = False):
    """X_i = A_i^{-1} B_i (custom_solve shape): returns [(X, info)].  mixed=True: FP32 LU + FP64 iterative
    refinement (srcwa_zgesv_batched_mixed), returning [(X, info, (fell_back, corrections, residual))]."""
    ctx = ctx or _lib.default_context()
    n = np.array([a.shape[0] for a in As], dtype=np.int32)
    nrhs = np.array([b.shape[1] for b in Bs], dtype=np.int32)
    A, B = _pack(As), _pack(Bs)
    X = np.zeros_like(B); info = np.zeros(len(As), dtype=np.int32)
    lib = _lib.load()
    if mixed:
        ref = np.zeros(3 * len(As))
        _lib.check(lib.srcwa_zgesv_batched_mixed(ctx.handle, len(As), _lib.ptr(n), _lib.ptr(nrhs), _lib.ptr(A),
                                                 _lib.ptr(B), _lib.ptr(X), _lib.ptr(info), _lib.ptr(ref)))
    else:
        _lib.check(lib.srcwa_zgesv_batched(ctx.handle, len(As), _lib.ptr(n), _lib.ptr(nrhs), _lib.ptr(A),
                                           _lib.ptr(B), _lib.ptr(X), _lib.ptr(info)))
    Xc = X.view(np.complex128)
    out, a = [], 0
    for i in range(len(As)):
        k, r = int(n[i]), int(nrhs[i])
        item = (Xc[a:a + k * r].reshape((k, r), order="F").copy(), int(info[i]))
        if mixed:
            item = item + ((int(ref[3 * i]), int(ref[3 * i + 1]), float(ref[3 * i + 2])),)
        out.append(item)
        a += k * r
    return out


def zgemm_batched(As, Bs, ctx: _lib.Context | None = None):
    """C_i = A_i B_i (custom_multiply shape)."""
    ctx = ctx or _lib.default_context()
    m = np.array([a.shape[0] for a in As], dtype=np.int32)
    k = np.array([a.shape[1] for a in As], dtype=np.int32)
    n = np.array([b.shape[1] for b in Bs], dtype=np.int32)
    A, B = _pack(As), _pack(Bs)
    C = np.zeros(2 * int(np.sum(m.astype(np.int64) * n)))
    _lib.check(_lib.load().srcwa_zgemm_batched(ctx.handle, len(As), _lib.ptr(m), _lib.ptr(n), _lib.ptr(k),
                                               _lib.ptr(A), _lib.ptr(B), _lib.ptr(C)))
    Cc = C.view(np.complex128)
    out, a = [], 0
    for i in range(len(As)):
        out.append(Cc[a:a + m[i] * n[i]].reshape((m[i], n[i]), order="F").copy())
        a += int(m[i] * n[i])
    return out
