"""numpy restatement of the reference's planar (TE/TM) RCWA solve. Test infrastructure only.

Follows include/stablercwa/rcwa1d.hpp of the reference:
  * analytic Fourier coefficients          rcwa1d.hpp:366-388  (``fourier_coefficients``)
  * Toeplitz convolution matrix            rcwa1d.hpp:391-407  (``convolution_matrix``)
  * transverse wavevectors                 rcwa1d.hpp:410-415
  * layer matrix TE / TM (inverse rule)    rcwa1d.hpp:432-468  (``assemble_layer``)
  * branch / analytic modes                rcwa1d.hpp:514-544
  * zero / diagonal / dense operator algebra  rcwa1d.hpp:549-629 (``Op``)
  * interface / top / propagation blocks   rcwa1d.hpp:644-784
  * seam recursion ``stack_on``            rcwa1d.hpp:794-808
  * bottom-up recursion, top block, downward sweep, efficiencies   rcwa1d.hpp:963-1055
The eigensolve is numpy's LAPACK zgeev instead of CORTH+COMQR2: efficiencies do not depend on the
order or phase of the eigenpairs (the reference's own test_rcwa1d.cpp:885-898 shuffles and rescales
them), so any backward-stable eigensolver gives the same efficiencies to rounding. FFT coefficient
layers follow fourier.hpp:26-83 (``resolve_fft_samples``, ``fourier_coefficients_fft``) with numpy's FFT
in place of the vendored pocketfft (both compute the same DFT).
"""
from __future__ import annotations

import math

import numpy as np

from .flat import decode1d

PI = 3.141592653589793238462643383279502884


class NumericalFailure(Exception):
    pass


def fourier_coefficients(edges, values, period, hmax):
    """c_h, h = -hmax..hmax, of a piecewise-constant profile (rcwa1d.hpp:366-388)."""
    c = np.zeros(2 * hmax + 1, dtype=np.complex128)
    e = np.asarray(edges, dtype=np.float64)
    v = np.asarray(values, dtype=np.complex128)
    for h in range(-hmax, hmax + 1):
        if h == 0:
            c[hmax] = np.sum(v * (e[1:] - e[:-1]) / period)
        else:
            arg = -2.0 * PI * h * e / period
            ph = np.cos(arg) + 1j * np.sin(arg)
            w = -2.0 * PI * h
            t = v * (ph[1:] - ph[:-1])
            c[h + hmax] = np.sum(t.imag / w - 1j * (t.real / w))
    return c


MAX_FFT_SAMPLES = 4194304


def resolve_fft_samples(M, samples=0):
    """The FFT sample count (fourier.hpp:26-44): fft_samples, or the smallest power of two >= max(8192,
    64 (2M+1)); more than 4M samples are needed."""
    if M > (MAX_FFT_SAMPLES - 1) // 4:
        raise ValueError("rcwa: FFT truncation requires more than 4194304 samples")
    if samples == 0:
        target = max(8192, 64 * (2 * M + 1))
        if target > MAX_FFT_SAMPLES:
            raise ValueError("rcwa: automatic FFT exceeds 4194304 samples")
        samples = 1
        while samples < target:
            samples *= 2
    if samples <= 4 * M:
        raise ValueError("rcwa: fft_samples must be at least 4*M+1")
    return samples


def fourier_coefficients_fft(edges, values, period, hmax, samples):
    """c_h, h = -hmax..hmax, from samples at the cell midpoints (j + 1/2) period / N: a forward DFT / N,
    then the phase of the origin and of the half cell (fourier.hpp:46-83)."""
    e = np.asarray(edges, dtype=np.float64)
    v = np.asarray(values, dtype=np.complex128)
    j = np.arange(samples, dtype=np.float64)
    u = (j + 0.5) / samples * period
    seg = np.zeros(samples, dtype=np.int64)
    for k in range(1, len(v)):                       # the reference's walk: u >= edges[k] - edges[0]
        seg += (u >= e[k] - e[0]).astype(np.int64)
    spectrum = np.fft.fft(v[seg]) / samples
    origin = math.fmod(e[0], period)
    if origin < 0:
        origin += period
    offset = origin / period + 0.5 / samples
    c = np.zeros(2 * hmax + 1, dtype=np.complex128)
    for h in range(-hmax, hmax + 1):
        b = samples + h if h < 0 else h
        arg = -2.0 * PI * h * offset
        c[h + hmax] = spectrum[b] * (math.cos(arg) + 1j * math.sin(arg))
    return c


def convolution_matrix(edges, values, period, M, method=0, samples=0):
    """[[f]](m, p) = c_{m-p} (rcwa1d.hpp:391-407), by the layer's Fourier method."""
    if method == 1:
        if not np.all(np.isfinite(np.asarray(values, dtype=np.complex128))):
            raise NumericalFailure()
        c = fourier_coefficients_fft(edges, values, period, 2 * M, resolve_fft_samples(M, samples))
    else:
        c = fourier_coefficients(edges, values, period, 2 * M)
    n = 2 * M + 1
    idx = np.arange(n)[:, None] - np.arange(n)[None, :] + 2 * M
    return c[idx]


def transverse_wavevectors(nI, theta_deg, wavelength, period, M):
    s = nI * math.sin(theta_deg * PI / 180.0)
    return np.array([s - (j - M) * wavelength / period for j in range(2 * M + 1)], dtype=np.float64)


def _solve(a, b):
    """Partial-pivot LU solve; an exactly singular matrix is a numerical failure (rcwa1d.hpp:248-281)."""
    try:
        x = np.linalg.solve(a, b)
    except np.linalg.LinAlgError as exc:
        raise NumericalFailure() from exc
    return x


def assemble_layer(layer, period, kx, M, pol):
    """Layer matrix A and [[1/eps]] (TM) (rcwa1d.hpp:432-468)."""
    edges = [0.0, period] if layer.homogeneous else layer.edges
    vals = [layer.n] if layer.homogeneous else layer.values
    eps = [v * v for v in vals]
    inv = [1.0 / (v * v) for v in vals]
    method = 0 if layer.homogeneous else layer.fourier_method
    e = convolution_matrix(edges, eps, period, M, method, layer.fft_samples)
    if pol == 0:
        a = -e
        a[np.diag_indices_from(a)] += kx * kx
        return a, None
    inv_eps = convolution_matrix(edges, inv, period, M, method, layer.fft_samples)
    b = _solve(e, np.diag(kx.astype(np.complex128))) * kx[:, None]
    b[np.diag_indices_from(b)] -= 1.0
    return _solve(inv_eps, b), inv_eps


def branch(q):
    """Re q >= 0; propagating (Re q ~ 0) takes Im q <= 0 (rcwa1d.hpp:514-518)."""
    q = np.where(q.real < 0.0, -q, q)
    q = np.where((q.real < 1e-10) & (q.imag > 0.0), -q, q)
    return q


class Modes:
    def __init__(self):
        self.analytic = False
        self.finite_zero_basis = False
        self.q = None
        self.p = None
        self.W = None
        self.V = None


def analytic_modes(n, kx, pol, finite_zero_basis=False):
    """Homogeneous medium: q = -i kz, p = q (TE) or q/n^2 (TM) (rcwa1d.hpp:522-544)."""
    m = Modes()
    m.analytic = True
    m.finite_zero_basis = finite_zero_basis
    kz = np.sqrt(n * n - kx.astype(np.complex128) ** 2)
    kz = np.where(kz.imag < 0.0, -kz, kz)
    kz = np.where((np.abs(kz.imag) < 1e-12) & (kz.real < 0.0), -kz, kz)
    m.q = kz.imag - 1j * kz.real
    basis = np.where((m.q == 0) & finite_zero_basis, 1.0 + 0j, m.q)
    m.p = basis if pol == 0 else basis / (n * n)
    return m


# Operators: ('zero', None) | ('diag', vector) | ('dense', matrix)   (rcwa1d.hpp:549-629)
ZERO = ("zero", None)


def op_mul(a, b):
    if a[0] == "zero" or b[0] == "zero":
        return ZERO
    if a[0] == "diag" and b[0] == "diag":
        return ("diag", a[1] * b[1])
    if a[0] == "diag":
        return ("dense", a[1][:, None] * b[1])
    if b[0] == "diag":
        return ("dense", a[1] * b[1][None, :])
    return ("dense", a[1] @ b[1])


def op_add(a, b):
    if b[0] == "zero":
        return a
    if a[0] == "zero":
        return b
    if a[0] == "diag" and b[0] == "diag":
        return ("diag", a[1] + b[1])
    if a[0] == "dense" and b[0] == "dense":
        return ("dense", a[1] + b[1])
    if a[0] == "dense":
        m = a[1].copy(); m[np.diag_indices_from(m)] += b[1]
        return ("dense", m)
    m = b[1].copy(); m[np.diag_indices_from(m)] += a[1]
    return ("dense", m)


def op_apply(a, x):
    if a[0] == "dense":
        return a[1] @ x
    if a[0] == "diag":
        return a[1] * x
    return np.zeros_like(x)


def op_to_dense(a, n):
    if a[0] == "dense":
        return a[1]
    if a[0] == "diag":
        return np.diag(a[1])
    return np.zeros((n, n), dtype=np.complex128)


def diagonal_interface(up, low):
    """Both sides homogeneous: per-order Fresnel blocks (rcwa1d.hpp:649-669)."""
    p1, p2 = up.p, low.p
    inv = 1.0 / (p1 + p2)
    return (("diag", 2.0 * p1 * inv), ("diag", (p2 - p1) * inv), ("diag", (p1 - p2) * inv), ("diag", 2.0 * p2 * inv))


def interface_matching(up, low):
    """T+ and T- (rcwa1d.hpp:672-692)."""
    n = up.q.size
    if low.analytic:
        mm = up.W
        nn = up.V / low.p[:, None]
    elif up.analytic:
        mm = _solve(low.W, np.eye(n, dtype=np.complex128))
        nn = _solve(low.V, np.diag(up.p))
    else:
        mm = _solve(low.W, up.W)
        nn = _solve(low.V, up.V)
    return 0.5 * (mm + nn), 0.5 * (mm - nn)


def interface_block(up, low):
    """(Tdd, Rdu, Rud, Tuu) of one interface (rcwa1d.hpp:694-712)."""
    if up.analytic and low.analytic:
        return diagonal_interface(up, low)
    n = up.q.size
    tp, tm = interface_matching(up, low)
    tpi = _solve(tp, np.eye(n, dtype=np.complex128))
    rdu = tm @ tpi
    tdd = tp - rdu @ tm
    rud = -(tpi @ tm)
    return (("dense", tdd), ("dense", rdu), ("dense", rud), ("dense", tpi))


def propagation_block(q, d, k0, finite_zero_basis):
    """exp(-k0 q d); exact zero roots of finite layers use 1/(1+h/2), (h/2)/(1+h/2) (rcwa1d.hpp:751-784)."""
    x = np.empty_like(q)
    refl = np.zeros_like(q)
    has_zero = False
    for j in range(q.size):
        if finite_zero_basis and q[j] == 0:
            h = k0 * d
            if not math.isfinite(h):
                raise NumericalFailure()
            den = 1.0 + 0.5 * h
            x[j] = 1.0 / den
            refl[j] = (0.5 * h) / den
            has_zero = True
        else:
            x[j] = np.exp(-k0 * d * q[j])
    if has_zero:
        return (("diag", x), ("diag", refl), ("diag", refl.copy()), ("diag", x.copy()))
    return (("diag", x), ZERO, ZERO, ("diag", x.copy()))


def stack_on(block, r, n):
    """g = (I - Rdu r)^-1 Tdd, r' = Rud + Tuu r g (rcwa1d.hpp:794-808)."""
    tdd, rdu, rud, tuu = block
    k = op_mul(rdu, r)
    if k[0] == "zero":
        g = tdd
    elif k[0] == "diag":
        g = op_mul(("diag", 1.0 / (1.0 - k[1])), tdd)
    else:
        g = ("dense", _solve(np.eye(n) - k[1], op_to_dense(tdd, n)))
    r2 = op_add(rud, op_mul(tuu, op_mul(r, g)))
    return g, r2


def solve(stack_flat, wavelength, theta_deg, M, pol=0):
    """One planar solve. Returns dict(R, T, DEr, DEt, eig_info). Analytic Fourier layers, fast path."""
    st = decode1d(stack_flat)
    n = 2 * M + 1
    k0 = 2.0 * PI / wavelength
    kx = transverse_wavevectors(st.nI.real, theta_deg, wavelength, st.period, M)
    L = len(st.layers)
    nan = np.full(n, np.nan)
    try:
        modes = [None] * (L + 2)
        modes[0] = analytic_modes(st.nI, kx, pol)
        modes[-1] = analytic_modes(st.nII, kx, pol)
        pure_films = all(l.homogeneous for l in st.layers)
        for j, layer in enumerate(st.layers):
            if layer.homogeneous:
                modes[j + 1] = analytic_modes(layer.n, kx, pol, True)
                continue
            a, inv_eps = assemble_layer(layer, st.period, kx, M, pol)
            if not np.all(np.isfinite(a)):
                raise NumericalFailure()
            lam, W = np.linalg.eig(a)
            W = W / np.linalg.norm(W, axis=0)[None, :]
            m = Modes()
            m.q = branch(np.sqrt(lam))
            m.W = W
            m.finite_zero_basis = True
            basis = np.where(m.q == 0, 1.0 + 0j, m.q)
            m.V = (W if pol == 0 else inv_eps @ W) * basis[None, :]
            modes[j + 1] = m
        nb = 2 * L + 1
        g = [None] * nb
        r = ZERO
        for b in range(nb - 1, 0, -1):
            j = b // 2
            if b % 2 == 0:
                block = interface_block(modes[j], modes[j + 1])
            else:
                block = propagation_block(modes[j + 1].q, st.layers[j].d, k0, modes[j + 1].finite_zero_basis)
            g[b], r = stack_on(block, r, n)
        # top block: only column M of Tdd and Rud (rcwa1d.hpp:715-746)
        up, low = modes[0], modes[1]
        if up.analytic and low.analytic:
            tdd_b, rdu_b, rud_b, tuu_b = diagonal_interface(up, low)
            tdd = np.zeros(n, dtype=np.complex128); tdd[M] = tdd_b[1][M]
            rud = np.zeros(n, dtype=np.complex128); rud[M] = rud_b[1][M]
            Rdu, Tuu = rdu_b, tuu_b
        else:
            tp, tm = interface_matching(up, low)
            tpi = _solve(tp, np.eye(n, dtype=np.complex128))
            rdu = tm @ tpi
            tm_e = tm[:, M]
            tdd = tp[:, M] - rdu @ tm_e
            rud = -(tpi @ tm_e)
            Rdu, Tuu = ("dense", rdu), ("dense", tpi)
        k = op_mul(Rdu, r)
        down = tdd.copy()
        if k[0] == "diag":
            down = down / (1.0 - k[1])
        elif k[0] == "dense":
            down = _solve(np.eye(n) - k[1], down)
        R = rud + op_apply(Tuu, op_apply(r, down))
        for b in range(1, nb):
            down = op_apply(g[b], down)
        T = down
        if pure_films:
            for j in range(n):
                if j == M:
                    continue
                if not np.isfinite(R[j]):
                    R[j] = 0
                if not np.isfinite(T[j]):
                    T[j] = 0
        if not (np.all(np.isfinite(R)) and np.all(np.isfinite(T))):
            raise NumericalFailure()
        qI, qII = modes[0].q, modes[-1].q
        den = complex(-k0 * qI[M].imag, k0 * qI[M].real)
        kIz = -k0 * qI.imag + 1j * k0 * qI.real
        kIIz = -k0 * qII.imag + 1j * k0 * qII.real
        DEr = np.abs(R) ** 2 * (kIz / den).real
        if pol == 0:
            factor = kIIz / den
        else:
            factor = (kIIz / (st.nII * st.nII)) / (den / (st.nI * st.nI))
        DEt = np.abs(T) ** 2 * factor.real
        if not (np.all(np.isfinite(DEr)) and np.all(np.isfinite(DEt))):
            raise NumericalFailure()
        return dict(R=R, T=T, DEr=DEr, DEt=DEt, eig_info=0)
    except NumericalFailure:
        cn = np.full(n, complex(np.nan, np.nan))
        return dict(R=cn, T=cn.copy(), DEr=nan, DEt=nan.copy(), eig_info=0)


def layer_matrix(stack_flat, layer_index, wavelength, theta_deg, M, pol=0):
    """The matrix the eigensolver receives for one layer (rcwa1d.hpp:1098-1110)."""
    st = decode1d(stack_flat)
    kx = transverse_wavevectors(st.nI.real, theta_deg, wavelength, st.period, M)
    return assemble_layer(st.layers[layer_index], st.period, kx, M, pol)[0]
