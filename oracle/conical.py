"""numpy restatement of the reference's conical solve. Test infrastructure only.

Follows include/stablercwa/rcwa1d_conical.hpp of the reference:
  * longitudinal / longitudinal0 / analytic_region   rcwa1d_conical.hpp:193-291
  * uniform-layer test                               rcwa1d_conical.hpp:293-299
  * assemble ([[eps]]^-1, C, B C, A)                  rcwa1d_conical.hpp:302-336
  * general_region (2N modes in the (s, t) basis)    rcwa1d_conical.hpp:341-411
  * triangular_root / triangular_exp / schur_region  rcwa1d_conical.hpp:413-502
  * diagonal / general interfaces, top block         rcwa1d_conical.hpp:507-795
  * solve_conical_impl                               rcwa1d_conical.hpp:678-982
The N x N eigensolves use numpy's LAPACK zgeev (efficiencies and Jones matrices do not depend on
eigenvector order or phase); the Schur-basis path uses scipy's complex Schur form in place of
CORTH + COMQR2 with keep_transform.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg

from .flat import decode1d
from .rcwa1d import (PI, ZERO, NumericalFailure, _solve, branch, convolution_matrix, op_add, op_mul, op_to_dense,
                     stack_on)

DEGENERATE_WARN = 1e-2


def longitudinal(n, kx, ky):
    z = n * n - kx * kx - ky * ky
    if z.imag == 0.0:
        return complex(math.sqrt(z.real), 0.0) if z.real >= 0.0 else complex(0.0, math.sqrt(-z.real))
    kz = complex(np.sqrt(complex(z)))
    if kz.imag < 0.0:
        kz = -kz
    if abs(kz.imag) < 1e-12 and kz.real < 0.0:
        kz = -kz
    return kz


def longitudinal0(n, nI, cosine):
    z = (n - nI) * (n + nI) + (nI * cosine) * (nI * cosine)
    if z.imag == 0.0:
        return complex(math.sqrt(z.real), 0.0) if z.real >= 0.0 else complex(0.0, math.sqrt(-z.real))
    kz = complex(np.sqrt(complex(z)))
    if kz.imag < 0.0:
        kz = -kz
    if abs(kz.imag) < 1e-12 and kz.real < 0.0:
        kz = -kz
    return kz


class Region:
    def __init__(self):
        self.analytic = False
        self.q = self.w = self.v = None
        self.W = self.V = None
        self.G = None


def analytic_region(n, kx, ky, nI, cosine):
    N = kx.size
    r = Region()
    r.analytic = True
    r.q = np.zeros(2 * N, complex); r.w = np.zeros(2 * N, complex); r.v = np.zeros(2 * N, complex)
    for j in range(N):
        kz = longitudinal0(n, nI, cosine) if 2 * j + 1 == N else longitudinal(n, kx[j], ky)
        q = complex(kz.imag, -kz.real)
        r.q[j] = r.q[N + j] = q
        r.w[j] = 1.0; r.v[j] = q
        r.w[N + j] = kz / n
        r.v[N + j] = complex(n.imag, -n.real)
    return r


def is_uniform(layer):
    return layer.homogeneous or all(v == layer.values[0] for v in layer.values)


def assemble(layer, period, kx, M):
    edges = [0.0, period] if layer.homogeneous else layer.edges
    vals = [layer.n] if layer.homogeneous else layer.values
    method = 0 if layer.homogeneous else layer.fourier_method
    e = convolution_matrix(edges, [v * v for v in vals], period, M, method, layer.fft_samples)
    N = 2 * M + 1
    inv_eps = _solve(e, np.eye(N, dtype=complex))
    c = _solve(convolution_matrix(edges, [1.0 / (v * v) for v in vals], period, M, method, layer.fft_samples),
               np.eye(N, dtype=complex))
    b = (kx[:, None] * inv_eps) * kx[None, :]
    b[np.diag_indices(N)] -= 1.0
    bc = b @ c
    a = -e
    a[np.diag_indices(N)] += kx * kx
    return dict(a=a, bc=bc, inv_eps=inv_eps, c=c)


def _eig(A):
    lam, W = np.linalg.eig(A)
    return lam, W / np.linalg.norm(W, axis=0)[None, :]


def general_region(te, tm, op, kx, ky, cs, sn):
    N = kx.size
    lam1, W1 = te
    lam2, W2 = tm
    r = Region()
    r.q = np.concatenate([branch(np.sqrt(lam1 + ky * ky)), branch(np.sqrt(lam2 + ky * ky))])
    cw = op["c"] @ W2
    coupling = op["inv_eps"] @ ((ky * kx)[:, None] * cw) if ky != 0.0 else np.zeros_like(cw)
    W = np.zeros((2 * N, 2 * N), complex); V = np.zeros((2 * N, 2 * N), complex)
    q1 = r.q[:N]; q2 = r.q[N:]
    ey = W1 * (q1 * q1)[None, :]
    hx = W1 * (-(lam1 * q1))[None, :]
    hy = (kx[:, None] * W1) * (-ky * q1)[None, :]
    W[:N, :N] = cs[:, None] * ey
    W[N:, :N] = sn[:, None] * ey
    V[:N, :N] = -(cs[:, None] * hx) - sn[:, None] * hy
    V[N:, :N] = cs[:, None] * hy - sn[:, None] * hx
    ex = W2 * lam2[None, :]
    ey2 = coupling
    hy2 = -(cw * q2[None, :])
    W[:N, N:] = cs[:, None] * ey2 - sn[:, None] * ex
    W[N:, N:] = cs[:, None] * ex + sn[:, None] * ey2
    V[:N, N:] = -(sn[:, None] * hy2)
    V[N:, N:] = cs[:, None] * hy2
    scale = np.sqrt(np.sum(np.abs(W) ** 2 + np.abs(V) ** 2, axis=0))
    if not np.all(np.isfinite(scale)) or np.any(scale <= 0):
        raise NumericalFailure()
    r.W = W / scale[None, :]
    r.V = V / scale[None, :]
    return r


def triangular_root(t):
    n = t.shape[0]
    g = np.zeros_like(t)
    for i in range(n):
        g[i, i] = branch(np.array([np.sqrt(t[i, i])]))[0]
        if g[i, i] == 0 or not np.isfinite(g[i, i]):
            raise NumericalFailure()
    for gap in range(1, n):
        for i in range(n - gap):
            j = i + gap
            v = t[i, j] - np.dot(g[i, i + 1:j], g[i + 1:j, j])
            den = g[i, i] + g[j, j]
            if den == 0:
                raise NumericalFailure()
            g[i, j] = v / den
    res = g @ g - t
    if np.linalg.norm(res, 1) > 1e-10 * max(np.linalg.norm(t, 1), np.finfo(float).tiny):
        raise NumericalFailure()
    return g


def triangular_exp(a):
    n = a.shape[0]
    size = np.linalg.norm(a, 1)
    scaling = max(0, int(math.ceil(math.log2(size) + 2.0))) if size > 0.25 else 0
    if scaling > 1020:
        raise NumericalFailure()
    b = a * math.ldexp(1.0, -scaling)
    e = np.eye(n, dtype=complex); term = e.copy()
    converged = n == 0
    for k in range(1, 41):
        if converged:
            break
        term = (term @ b) * (1.0 / k)
        e = e + term
        converged = np.linalg.norm(term, 1) <= np.finfo(float).eps * max(1.0, np.linalg.norm(e, 1))
    if not converged:
        raise NumericalFailure()
    d = np.diag(a)
    e[np.diag_indices(n)] = np.exp(d * math.ldexp(1.0, -scaling))
    for k in range(scaling):
        e = e @ e
        e[np.diag_indices(n)] = np.exp(d * math.ldexp(1.0, k + 1 - scaling))
    return e


def schur_region(op, kx, ky, cs, sn):
    N = kx.size
    n = 2 * N
    ie, A, C = op["inv_eps"], op["a"], op["c"]
    I = np.eye(N)
    p = np.zeros((n, n), complex); q = np.zeros((n, n), complex)
    p[:N, :N] = ky * kx[:, None] * ie
    p[:N, N:] = I - kx[:, None] * ie * kx[None, :]
    p[N:, :N] = ky * ky * ie - I
    p[N:, N:] = -ky * ie * kx[None, :]
    q[:N, :N] = np.diag(ky * kx); q[:N, N:] = -A
    q[N:, :N] = ky * ky * I - C; q[N:, N:] = -np.diag(ky * kx)
    re = np.zeros((n, n)); rv = np.zeros((n, n))
    for i in range(N):
        re[i, i] = -sn[i]; re[i, N + i] = re[N + i, i] = cs[i]; re[N + i, N + i] = sn[i]
        rv[i, i] = -cs[i]; rv[i, N + i] = rv[N + i, i] = -sn[i]; rv[N + i, N + i] = cs[i]
    p = re @ p @ rv
    q = rv @ q @ re
    h = p @ q
    T, Z = scipy.linalg.schur(h, output="complex")
    r = Region()
    r.G = triangular_root(T)
    r.q = np.diag(r.G).copy()
    r.W = Z
    r.V = scipy.linalg.solve_triangular(r.G.T, (q @ Z).T, lower=True).T
    return r


def diagonal_interface(up, low):
    w1, v1, w2, v2 = up.w, up.v, low.w, low.v
    a = w1 * v2; b = v1 * w2
    inv = 1.0 / (a + b)
    return (("diag", 2.0 * (w1 * v1) * inv), ("diag", (a - b) * inv), ("diag", (b - a) * inv),
            ("diag", 2.0 * (w2 * v2) * inv))


def interface_matching(up, low):
    n = up.q.size
    if low.analytic:
        if np.any(low.w == 0) or np.any(low.v == 0):
            raise NumericalFailure()
        mm = up.W / low.w[:, None]
        nn = up.V / low.v[:, None]
    elif up.analytic:
        mm = _solve(low.W, np.diag(up.w))
        nn = _solve(low.V, np.diag(up.v))
    else:
        mm = _solve(low.W, up.W)
        nn = _solve(low.V, up.V)
    return 0.5 * (mm + nn), 0.5 * (mm - nn)


def interface_block(up, low):
    if up.analytic and low.analytic:
        return diagonal_interface(up, low)
    n = up.q.size
    tp, tm = interface_matching(up, low)
    tpi = _solve(tp, np.eye(n, dtype=complex))
    rdu = tm @ tpi
    return (("dense", tp - rdu @ tm), ("dense", rdu), ("dense", -(tpi @ tm)), ("dense", tpi))


def _apply(a, x):
    if a[0] == "dense":
        return a[1] @ x
    if a[0] == "diag":
        return a[1][:, None] * x
    return np.zeros_like(x)


def solve(stack_flat, wavelength, theta_deg, phi_deg, M):
    """One conical solve; returns dict(Jr, Jt (N x 4: ss, sp, ps, pp), DEr, DEt, flux_r, flux_t, n_schur)."""
    st = decode1d(stack_flat)
    N = 2 * M + 1
    n = 2 * N
    layers = [l for l in st.layers if l.d > 0.0]
    L = len(layers)
    k0 = 2.0 * PI / wavelength
    theta = theta_deg * PI / 180.0; phi = phi_deg * PI / 180.0
    nI = st.nI.real
    sine = nI * math.sin(theta)
    ky = sine * math.sin(phi)
    flux_unit = nI * math.cos(theta)
    kx = np.zeros(N); cs = np.zeros(N); sn = np.zeros(N)
    fr = np.zeros(N); ft = np.zeros(N)
    for j in range(N):
        kxj = sine * math.cos(phi) - (j - M) * wavelength / st.period
        kt = math.hypot(kxj, ky)
        kx[j] = kxj
        if kt <= 1e-14:
            cs[j], sn[j] = math.cos(phi), math.sin(phi)
        else:
            cs[j], sn[j] = kxj / kt, ky / kt
        kzI = longitudinal0(st.nI, nI, math.cos(theta)) if j == M else longitudinal(st.nI, kxj, ky)
        kzII = longitudinal0(st.nII, nI, math.cos(theta)) if j == M else longitudinal(st.nII, kxj, ky)
        fr[j] = kzI.real / flux_unit
        ft[j] = kzII.real / flux_unit
    nanJ = np.full((N, 4), complex(np.nan, np.nan))
    n_schur = 0
    try:
        modes = [None] * (L + 2)
        modes[0] = analytic_region(st.nI, kx, ky, nI, math.cos(theta))
        modes[-1] = analytic_region(st.nII, kx, ky, nI, math.cos(theta))
        pure = all(is_uniform(l) for l in layers)
        ky_scale = abs(ky) * math.sqrt(abs(ky))
        for j, layer in enumerate(layers):
            if is_uniform(layer):
                idx = layer.n if layer.homogeneous else layer.values[0]
                modes[j + 1] = analytic_region(idx, kx, ky, nI, math.cos(theta))
                for k in range(N):
                    if modes[j + 1].q[k] == 0 and (not pure or k == M):
                        raise NumericalFailure()
                continue
            op = assemble(layer, st.period, kx, M)
            if not (np.all(np.isfinite(op["a"])) and np.all(np.isfinite(op["bc"]))):
                raise NumericalFailure()
            te, tm = _eig(op["a"]), _eig(op["bc"])
            local_min = min(np.min(np.abs(te[0])), np.min(np.abs(tm[0])))
            if ky != 0.0 and local_min < DEGENERATE_WARN * ky_scale:
                n_schur += 1
                modes[j + 1] = schur_region(op, kx, ky, cs, sn)
            else:
                modes[j + 1] = general_region(te, tm, op, kx, ky, cs, sn)
        nb = 2 * L + 1
        g = [None] * nb
        r = ZERO
        for b in range(nb - 1, 0, -1):
            j = b // 2
            if b % 2 == 0:
                blk = interface_block(modes[j], modes[j + 1])
            else:
                m = modes[j + 1]
                if m.G is not None:
                    x = triangular_exp(m.G * (-k0 * layers[j].d))
                    blk = (("dense", x), ZERO, ZERO, ("dense", x))
                else:
                    x = np.exp(-k0 * layers[j].d * m.q)
                    blk = (("diag", x), ZERO, ZERO, ("diag", x.copy()))
            g[b], r = stack_on(blk, r, n)
        cols = [M, N + M]
        up, low = modes[0], modes[1]
        if up.analytic and low.analytic:
            tdd_b, rdu_b, rud_b, tuu_b = diagonal_interface(up, low)
            tdd = np.zeros((n, 2), complex); rud = np.zeros((n, 2), complex)
            for k, c in enumerate(cols):
                tdd[c, k] = tdd_b[1][c]
                rud[c, k] = rud_b[1][c]
            Rdu, Tuu = rdu_b, tuu_b
        else:
            tp, tm = interface_matching(up, low)
            tpi = _solve(tp, np.eye(n, dtype=complex))
            rdu = tm @ tpi
            tm_e = tm[:, cols]
            tdd = tp[:, cols] - rdu @ tm_e
            rud = -(tpi @ tm_e)
            Rdu, Tuu = ("dense", rdu), ("dense", tpi)
        k = op_mul(Rdu, r)
        down = tdd.copy()
        if k[0] == "diag":
            down = down / (1.0 - k[1])[:, None]
        elif k[0] == "dense":
            down = _solve(np.eye(n) - k[1], down)
        upa = rud + _apply(Tuu, _apply(r, down))
        for b in range(1, nb):
            down = _apply(g[b], down)
        Jr = np.stack([upa[:N, 0], upa[:N, 1], -upa[N:, 0], -upa[N:, 1]], axis=1)
        Jt = np.stack([down[:N, 0], down[:N, 1], down[N:, 0], down[N:, 1]], axis=1)
        if pure:
            for j in range(N):
                if j == M:
                    continue
                Jr[j][~np.isfinite(Jr[j])] = 0
                Jt[j][~np.isfinite(Jt[j])] = 0
        if not (np.all(np.isfinite(Jr)) and np.all(np.isfinite(Jt))):
            raise NumericalFailure()
        DEr = fr[:, None] * np.abs(Jr) ** 2
        DEt = ft[:, None] * np.abs(Jt) ** 2
        return dict(Jr=Jr, Jt=Jt, DEr=DEr, DEt=DEt, flux_r=fr, flux_t=ft, n_schur=n_schur, ok=True)
    except NumericalFailure:
        return dict(Jr=nanJ, Jt=nanJ.copy(), DEr=np.full((N, 4), np.nan), DEt=np.full((N, 4), np.nan),
                    flux_r=fr, flux_t=ft, n_schur=n_schur, ok=False)
