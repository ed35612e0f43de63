"""ctypes loader of oracle/_ref/libsrcwa_ref.so: the reference's own solver built from its headers by
oracle/Makefile.  Test infrastructure only; absent wherever the reference tree was absent at build time."""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libsrcwa_ref.so")

_lib = None


def available() -> bool:
    return os.path.exists(LIB)


def load():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(LIB)
        _lib.ref_last_error.restype = ctypes.c_char_p
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def solve1d(lib, stack_flat, wavelength, theta, M, pol, fast_path=True):
    N = 2 * M + 1
    f = np.ascontiguousarray(stack_flat, dtype=np.float64)
    R = np.zeros(2 * N); T = np.zeros(2 * N); DEr = np.zeros(N); DEt = np.zeros(N)
    info = ctypes.c_int(); res = ctypes.c_double()
    rc = lib.ref_solve1d(_p(f), ctypes.c_double(wavelength), ctypes.c_double(theta), M, pol, int(fast_path),
                         _p(R), _p(T), _p(DEr), _p(DEt), ctypes.byref(info), ctypes.byref(res))
    if rc != 0:
        raise ValueError(lib.ref_last_error().decode())
    return dict(R=R.view(np.complex128), T=T.view(np.complex128), DEr=DEr, DEt=DEt, eig_info=info.value,
                max_eig_residual=res.value)


def sweep1d(lib, stack_flat, wavelengths, thetas, M, pol, threads):
    lib.ref_set_threads(int(threads))
    N = 2 * M + 1
    f = np.ascontiguousarray(stack_flat, dtype=np.float64)
    w = np.ascontiguousarray(wavelengths, dtype=np.float64)
    t = np.ascontiguousarray(thetas, dtype=np.float64)
    DEr = np.zeros(len(w) * N); DEt = np.zeros(len(w) * N)
    rc = lib.ref_sweep1d(_p(f), len(w), _p(w), _p(t), M, pol, _p(DEr), _p(DEt))
    if rc != 0:
        raise ValueError(lib.ref_last_error().decode())
    return DEr.reshape(len(w), N), DEt.reshape(len(w), N)


def eig(lib, A):
    n = A.shape[0]
    a = np.asfortranarray(A, dtype=np.complex128)
    flat = a.ravel(order="F").view(np.float64).copy()
    w = np.zeros(2 * n); V = np.zeros(2 * n * n); info = ctypes.c_int(); res = ctypes.c_double()
    rc = lib.ref_eig(n, _p(flat), _p(w), _p(V), ctypes.byref(info), ctypes.byref(res))
    if rc != 0:
        raise ValueError(lib.ref_last_error().decode())
    return w.view(np.complex128), V.view(np.complex128).reshape((n, n), order="F"), info.value, res.value


def solve1d_ex(lib, stack_flat, wavelength, theta, M, pol, max_iter=30, diagnostics=True):
    """ref_solve1d_ex: the reference's solve with an eigen iteration limit and Options::diagnostics;
    returns DEr, DEt and eig_info, n_eig, n_dense_solve, max_eig_residual, max_solve_backward_error."""
    N = 2 * M + 1
    f = np.ascontiguousarray(stack_flat, dtype=np.float64)
    R = np.zeros(2 * N); T = np.zeros(2 * N); DEr = np.zeros(N); DEt = np.zeros(N)
    counts = np.zeros(4, np.int32); met = np.zeros(2)
    rc = lib.ref_solve1d_ex(_p(f), ctypes.c_double(wavelength), ctypes.c_double(theta), M, pol, max_iter,
                            int(diagnostics), _p(R), _p(T), _p(DEr), _p(DEt), _p(counts), _p(met))
    if rc != 0:
        raise ValueError(lib.ref_last_error().decode())
    return dict(DEr=DEr, DEt=DEt, eig_info=int(counts[0]), n_eig=int(counts[1]), n_dense_solve=int(counts[2]),
                max_eig_residual=float(met[0]), max_solve_backward_error=float(met[1]))


def solve_conical_ex(lib, stack_flat, wavelength, theta, phi, M, max_iter=30, diagnostics=True):
    """ref_solve_conical_ex (as solve1d_ex, plus n_schur)."""
    N = 2 * M + 1
    f = np.ascontiguousarray(stack_flat, dtype=np.float64)
    Jr = np.zeros(8 * N); Jt = np.zeros(8 * N); DEr = np.zeros(4 * N); DEt = np.zeros(4 * N)
    counts = np.zeros(4, np.int32); met = np.zeros(2)
    rc = lib.ref_solve_conical_ex(_p(f), ctypes.c_double(wavelength), ctypes.c_double(theta), ctypes.c_double(phi),
                                  M, max_iter, int(diagnostics), _p(Jr), _p(Jt), _p(DEr), _p(DEt), _p(counts),
                                  _p(met))
    if rc != 0:
        raise ValueError(lib.ref_last_error().decode())
    return dict(DEr=DEr, DEt=DEt, eig_info=int(counts[0]), n_eig=int(counts[1]), n_dense_solve=int(counts[2]),
                n_schur=int(counts[3]), max_eig_residual=float(met[0]), max_solve_backward_error=float(met[1]))


def solve2d_ex(lib, stack_flat, wavelength, theta, phi, M1, M2, jones, max_iter=30, diagnostics=True):
    """ref_solve2d_ex: status, eig_info, n_eig, n_dense_solve, max_eig_residual."""
    N = (2 * M1 + 1) * (2 * M2 + 1)
    f = np.ascontiguousarray(stack_flat, dtype=np.float64)
    j = np.ascontiguousarray(jones, dtype=np.float64)
    DEr = np.zeros(N); DEt = np.zeros(N)
    Rs, Rp, Ts, Tp = (np.zeros(2 * N) for _ in range(4))
    st = ctypes.c_int(); counts = np.zeros(4, np.int32); met = np.zeros(2)
    rc = lib.ref_solve2d_ex(_p(f), ctypes.c_double(wavelength), ctypes.c_double(theta), ctypes.c_double(phi), M1,
                            M2, _p(j), max_iter, int(diagnostics), _p(DEr), _p(DEt), _p(Rs), _p(Rp), _p(Ts), _p(Tp),
                            ctypes.byref(st), _p(counts), _p(met))
    if rc != 0:
        raise ValueError(lib.ref_last_error().decode())
    return dict(DEr=DEr, DEt=DEt, status=st.value, eig_info=int(counts[0]), n_eig=int(counts[1]),
                n_dense_solve=int(counts[2]), max_eig_residual=float(met[0]))


def _stack_ptrs(flats):
    arrs = [np.ascontiguousarray(f, dtype=np.float64) for f in flats]
    ptrs = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return arrs, ptrs


def solve_conical_many(lib, flats, wavelengths, thetas, phis, Ms, threads):
    """ref_solve_conical_many on `threads` pool threads: per-instance total efficiency and eig_info."""
    lib.ref_set_threads(int(threads))
    n = len(flats)
    keep, ptrs = _stack_ptrs(flats)
    w, t, p = (np.ascontiguousarray(x, dtype=np.float64) for x in (wavelengths, thetas, phis))
    m = np.ascontiguousarray(Ms, dtype=np.int32)
    energy = np.zeros(n); info = np.zeros(n, np.int32)
    rc = lib.ref_solve_conical_many(n, ptrs, _p(w), _p(t), _p(p), _p(m), _p(energy), _p(info))
    if rc != 0:
        raise ValueError(lib.ref_last_error().decode())
    return energy, info


def solve2d_many(lib, flats, wavelengths, thetas, phis, M1s, M2s, jones, threads):
    """ref_solve2d_many on `threads` pool threads: per-instance total efficiency and status.
    jones: n x 4 doubles (Es.re, Es.im, Ep.re, Ep.im)."""
    lib.ref_set_threads(int(threads))
    n = len(flats)
    keep, ptrs = _stack_ptrs(flats)
    w, t, p = (np.ascontiguousarray(x, dtype=np.float64) for x in (wavelengths, thetas, phis))
    m1 = np.ascontiguousarray(M1s, dtype=np.int32); m2 = np.ascontiguousarray(M2s, dtype=np.int32)
    j = np.ascontiguousarray(jones, dtype=np.float64).reshape(n, 4)
    energy = np.zeros(n); status = np.zeros(n, np.int32)
    rc = lib.ref_solve2d_many(n, ptrs, _p(w), _p(t), _p(p), _p(m1), _p(m2), _p(j), _p(energy), _p(status))
    if rc != 0:
        raise ValueError(lib.ref_last_error().decode())
    return energy, status


def planar_fields(lib, stack_flat, wavelength, theta, M, pol, zs, nx=0, y=0.0):
    """rcwa::solve_with_fields + field_at (+ field_map with nx > 0) of the reference (rcwa1d_fields.hpp:55-206).
    Returns dict(available, harm [6, N, nz], grid [6, nz, nx] or None, DEr, DEt); harm/grid are None when the
    context is unavailable."""
    st = np.ascontiguousarray(stack_flat, dtype=np.float64)
    z = np.ascontiguousarray(np.atleast_1d(np.asarray(zs, dtype=np.float64)))
    N, nz = 2 * M + 1, z.size
    harm = np.zeros(12 * N * nz)
    grid = np.zeros(12 * max(nx, 0) * nz)
    DEr = np.zeros(N); DEt = np.zeros(N)
    avail = ctypes.c_int(0)
    rc = lib.ref_planar_fields(_p(st), ctypes.c_double(wavelength), ctypes.c_double(theta), int(M), int(pol), int(nz),
                               _p(z), int(nx), ctypes.c_double(y), ctypes.byref(avail), _p(harm), _p(grid), _p(DEr),
                               _p(DEt))
    if rc != 0:
        raise RuntimeError(lib.ref_last_error().decode())
    ok = bool(avail.value)
    h = harm.view(np.complex128).reshape(6, nz, N).transpose(0, 2, 1) if ok else None
    g = grid.view(np.complex128).reshape(6, nz, nx) if ok and nx > 0 else None
    return dict(available=ok, harm=h, grid=g, DEr=DEr, DEt=DEt)


def conical_fields(lib, stack_flat, wavelength, theta, phi, M, jones, zs, nx=0, y=0.0):
    """rcwa::solve_conical_with_fields + field_at (+ field_map with nx > 0) of the reference for the Jones vector
    jones = (Es, Ep) (rcwa1d_conical_fields.hpp:56-262).  Returns dict(available, n_schur, harm [6, N, nz],
    grid [6, nz, nx] or None)."""
    st = np.ascontiguousarray(stack_flat, dtype=np.float64)
    z = np.ascontiguousarray(np.atleast_1d(np.asarray(zs, dtype=np.float64)))
    j = np.array([complex(jones[0]).real, complex(jones[0]).imag, complex(jones[1]).real, complex(jones[1]).imag])
    N, nz = 2 * M + 1, z.size
    harm = np.zeros(12 * N * nz)
    grid = np.zeros(12 * max(nx, 0) * nz)
    avail = ctypes.c_int(0); ns = ctypes.c_int(0)
    rc = lib.ref_conical_fields(_p(st), ctypes.c_double(wavelength), ctypes.c_double(theta), ctypes.c_double(phi),
                                int(M), _p(j), int(nz), _p(z), int(nx), ctypes.c_double(y), ctypes.byref(avail),
                                ctypes.byref(ns), _p(harm), _p(grid))
    if rc != 0:
        raise RuntimeError(lib.ref_last_error().decode())
    ok = bool(avail.value)
    h = harm.view(np.complex128).reshape(6, nz, N).transpose(0, 2, 1) if ok else None
    g = grid.view(np.complex128).reshape(6, nz, nx) if ok and nx > 0 else None
    return dict(available=ok, n_schur=ns.value, harm=h, grid=g)
