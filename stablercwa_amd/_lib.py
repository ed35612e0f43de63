"""ctypes binding of libsrcwa_amd.so (include/srcwa.h).

The library is built in-tree by ``python -m stablercwa_amd.build``; a missing library is an error,
never a silent fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsrcwa_amd.so")

OK, EINVAL, EDEVICE, ENOMEM, EUNSUPPORTED = 0, -1, -2, -3, -4
EUNAVAILABLE, ENUMERICAL = -5, -6   # near fields: unavailable context (logic_error), non-finite values (runtime_error)
PLAN_DIAGNOSTICS = 1
PLAN_MIXED = 2          # star-product solves by FP32 LU + FP64 iterative refinement
PLAN_POINT_OPS = 4      # every point computes its own wavelength-invariant layer operators
PLAN_FIELDS = 8         # planar plans keep each solve's near-field capture (field_at / field_map)

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_i64_p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

# symbol -> (restype, argtypes)
PROTOTYPES = {
    "srcwa_last_error": (ctypes.c_char_p, []),
    "srcwa_version": (ctypes.c_char_p, []),
    "srcwa_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "srcwa_context_destroy": (None, [_vp]),
    "srcwa_context_sync": (ctypes.c_int, [_vp]),
    "srcwa_plan1d_create": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                           ctypes.c_int64, ctypes.POINTER(_vp)]),
    "srcwa_plan_conical_create": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp,
                                                 _vp, ctypes.c_int64, ctypes.POINTER(_vp)]),
    "srcwa_plan2d_create": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                           _vp, _vp, _vp, ctypes.c_int64, ctypes.POINTER(_vp)]),
    "srcwa_plan1d_bytes": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                          ctypes.c_int64, _vp]),
    "srcwa_plan_conical_bytes": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp,
                                                _vp, ctypes.c_int64, _vp]),
    "srcwa_plan2d_bytes": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                          _vp, _vp, _vp, ctypes.c_int64, _vp]),
    "srcwa_plan_execute": (ctypes.c_int, [_vp]),
    "srcwa_plan_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "srcwa_plan_set_max_iter": (ctypes.c_int, [_vp, ctypes.c_int]),
    "srcwa_plan_poll": (ctypes.c_int, [_vp]),
    "srcwa_pool_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.POINTER(_vp)]),
    "srcwa_pool_destroy": (ctypes.c_int, [_vp]),
    "srcwa_pool_free_bytes": (ctypes.c_int64, [_vp]),
    "srcwa_context_set_pool": (ctypes.c_int, [_vp, _vp]),
    "srcwa_context_set_plan_flags": (ctypes.c_int, [_vp, ctypes.c_int]),
    "srcwa_plan_diagnostics": (ctypes.c_int, [_vp, _vp]),
    "srcwa_plan_refinement": (ctypes.c_int, [_vp, _vp]),
    "srcwa_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(_vp)]),
    "srcwa_server_submit": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_int64)]),
    "srcwa_server_wait": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]),
    "srcwa_server_result": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "srcwa_server_release": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "srcwa_server_stats": (ctypes.c_int, [_vp, _vp]),
    "srcwa_server_last_error": (ctypes.c_char_p, []),
    "srcwa_server_destroy": (None, [_vp]),
    "srcwa_plan_sync": (ctypes.c_int, [_vp]),
    "srcwa_plan_results": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "srcwa_plan_orders": (ctypes.c_int64, [_vp]),
    "srcwa_plan1d_field_available": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "srcwa_plan1d_field_at": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "srcwa_plan1d_field_map": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_double,
                                              _vp]),
    "srcwa_plan_conical_field_at": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]),
    "srcwa_plan_conical_field_layers": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "srcwa_plan1d_field_layers": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "srcwa_plan_conical_field_map": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                                    ctypes.c_double, _vp, _vp]),
    "srcwa_plan_operator_sets": (ctypes.c_int, [_vp]),
    "srcwa_plan_geometry": (ctypes.c_int, [_vp, _vp]),
    "srcwa_plan_device_bytes": (ctypes.c_int64, [_vp]),
    "srcwa_plan_chunks": (ctypes.c_int, [_vp]),
    "srcwa_plan_stage_ms": (ctypes.c_int, [_vp, _vp]),
    "srcwa_plan_flops": (ctypes.c_int, [_vp, _vp]),
    "srcwa_plan_destroy": (None, [_vp]),
    "srcwa_solve1d_batch": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                           _vp, _vp, _vp, _vp, _vp, _vp]),
    "srcwa_zgeev_batched": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
    "srcwa_zgeev_batched_iter": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]),
    "srcwa_zgesv_batched": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "srcwa_zgesv_batched_mixed": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "srcwa_zgemm_batched": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "srcwa_zgeev_batched_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
}

_lib = None
_lock = threading.Lock()


def load() -> ctypes.CDLL:
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"{LIB_PATH} is not built; run `python -m stablercwa_amd.build`")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in PROTOTYPES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


class SrcwaError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class FieldUnavailable(SrcwaError):
    """A near-field context is unavailable (the reference's std::logic_error)."""


def check(code: int) -> None:
    if code != OK:
        msg = load().srcwa_last_error().decode("utf-8", "replace")
        if code == EINVAL:
            raise ValueError(msg)
        if code == EUNSUPPORTED:
            raise NotImplementedError(msg)
        if code == EUNAVAILABLE:
            raise FieldUnavailable(code, msg)
        raise SrcwaError(code, msg)


def ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


class Context:
    """A device context: one HIP stream on one GPU."""

    def __init__(self, device: int = 0, plan_flags: int = 0):
        lib = load()
        h = _vp()
        check(lib.srcwa_context_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device
        if plan_flags:
            check(lib.srcwa_context_set_plan_flags(h, int(plan_flags)))

    def close(self) -> None:
        if self.handle:
            load().srcwa_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    with _lock:
        pass
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx
