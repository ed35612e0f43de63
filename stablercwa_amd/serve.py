"""Python front end of the request server (``srcwa_server_*`` in include/srcwa.h).

    from stablercwa_amd.serve import Server
    with Server(hbm_gib=32, slots=4) as srv:
        rid = srv.submit_planar(stack, wavelengths, theta_deg=66.36, M=60, pol=0)
        res = srv.result(rid)            # waits; dict with R, T, DEr, DEt, status, eig_info, latency_s

Requests are queued on the native worker thread, which batches them continuously onto the GPU (up to
``slots`` batches in flight on separate HIP streams, admission by device-memory footprint).  Each
request's results are the same as solving it alone.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib


class _Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n_points", ctypes.c_int), ("stack", ctypes.c_void_p),
                ("stack_len", ctypes.c_int64), ("wavelength", ctypes.c_void_p), ("theta_deg", ctypes.c_void_p),
                ("phi_deg", ctypes.c_void_p), ("M", ctypes.c_int), ("M2", ctypes.c_int), ("pol", ctypes.c_int),
                ("factorization", ctypes.c_int), ("jones", ctypes.c_void_p)]


def _check(code: int) -> None:
    if code != _lib.OK:
        msg = _lib.load().srcwa_server_last_error().decode("utf-8", "replace")
        if code == _lib.EINVAL:
            raise ValueError(msg)
        if code == _lib.EUNSUPPORTED:
            raise NotImplementedError(msg)
        raise _lib.SrcwaError(code, msg)


class Server:
    def __init__(self, device: int = 0, hbm_gib: float = 32.0, slots: int = 4, max_batch_points: int = 4096):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _check(self.lib.srcwa_server_create(device, int(hbm_gib * (1 << 30)), slots, max_batch_points,
                                            ctypes.byref(h)))
        self.handle = h
        self._meta = {}

    # ---- submission
    def _submit(self, kind, stack_flat, wavelengths, thetas, phis=None, M=0, M2=0, pol=0, factorization=0,
                jones=None):
        st = np.ascontiguousarray(stack_flat, dtype=np.float64)
        wl = np.ascontiguousarray(np.atleast_1d(wavelengths), dtype=np.float64)
        P = wl.size
        th = np.ascontiguousarray(np.broadcast_to(np.asarray(thetas, dtype=np.float64), (P,)))
        ph = None if phis is None else np.ascontiguousarray(np.broadcast_to(np.asarray(phis, dtype=np.float64), (P,)))
        jv = None if jones is None else np.array([jones[0].real, jones[0].imag, jones[1].real, jones[1].imag])
        req = _Request(kind, P, st.ctypes.data, st.size, wl.ctypes.data, th.ctypes.data,
                       None if ph is None else ph.ctypes.data, M, M2, pol, factorization,
                       None if jv is None else jv.ctypes.data)
        rid = ctypes.c_int64()
        _check(self.lib.srcwa_server_submit(self.handle, ctypes.byref(req), ctypes.byref(rid)))
        n1 = 2 * M + 1
        N = n1 * (2 * M2 + 1) if kind == 3 else n1
        self._meta[rid.value] = (kind, P, N)
        return rid.value

    def submit_planar(self, stack, wavelengths, theta_deg, M: int, pol: int = 0) -> int:
        return self._submit(1, stack.flat(), wavelengths, theta_deg, M=M, pol=int(pol))

    def submit_conical(self, stack, wavelengths, theta_deg, phi_deg, M: int) -> int:
        return self._submit(2, stack.flat(), wavelengths, theta_deg, phi_deg, M=M)

    def submit_crossed(self, stack, wavelengths, theta_deg, phi_deg, M1: int, M2: int, jones,
                       factorization: str = "li") -> int:
        return self._submit(3, stack.flat(), wavelengths, theta_deg, phi_deg, M=M1, M2=M2,
                            factorization=0 if factorization == "li" else 1, jones=jones)

    # ---- completion
    def wait(self, rid: int, timeout_s: float = -1.0) -> bool:
        done = ctypes.c_int()
        _check(self.lib.srcwa_server_wait(self.handle, rid, float(timeout_s), ctypes.byref(done)))
        return bool(done.value)

    def result(self, rid: int, release: bool = True):
        self.wait(rid)
        kind, P, N = self._meta[rid]
        rw = 2 if kind == 1 else 8
        dw = 4 if kind == 2 else 1
        R = np.empty(rw * P * N); T = np.empty(rw * P * N)
        DEr = np.empty(dw * P * N); DEt = np.empty(dw * P * N)
        st = np.empty(P, dtype=np.int32); info = np.empty(P, dtype=np.int32)
        times = np.zeros(2)
        _check(self.lib.srcwa_server_result(self.handle, rid, _lib.ptr(R), _lib.ptr(T), _lib.ptr(DEr), _lib.ptr(DEt),
                                            _lib.ptr(st), _lib.ptr(info), _lib.ptr(times)))
        if release:
            self.release(rid)
        return dict(kind=kind, points=P, N=N, R=R.view(np.complex128), T=T.view(np.complex128), DEr=DEr, DEt=DEt,
                    status=st, eig_info=info, latency_s=float(times[0]), queued_s=float(times[1]))

    def release(self, rid: int) -> None:
        _check(self.lib.srcwa_server_release(self.handle, rid))
        self._meta.pop(rid, None)

    def stats(self):
        s = np.zeros(4, dtype=np.int64)
        _check(self.lib.srcwa_server_stats(self.handle, _lib.ptr(s)))
        return dict(batches=int(s[0]), max_inflight=int(s[1]), requests=int(s[2]), points=int(s[3]))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.srcwa_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
