"""numpy restatement of the reference's crossed-grating (2-D) solve. Test infrastructure only.

Follows the reference's include/stablercwa/rcwa2d_operators.hpp and rcwa2d.hpp:
  * fourier_1d / toeplitz_1d / seg_fourier / floor_mod / eps_at / merge_x_cuts   rcwa2d_operators.hpp:78-174
  * assemble_xfirst / assemble_yfirst                                          rcwa2d_operators.hpp:202-241
  * layer_operators (Li factorisation: Eps, Cx, Cy; Laurent: Cx = Cy = Eps)     rcwa2d_operators.hpp:349-429
  * kz_over_k0, branch, homogeneous_modes (2 x 2 block-diagonal V)             rcwa2d.hpp:380-404
  * structured_modes (Ei, P, Q, H = P Q, eig, V = Q W / q)                      rcwa2d.hpp:412-534
  * interface_bd, interface_matching, direct_dense, direct_bd_over_dense,
    stack_on (the three branches), top_block (direct form)                    rcwa2d.hpp:560-910
  * solve_impl: Rayleigh check, recursion, downward sweep, efficiencies, s/p    rcwa2d.hpp:1042-1339
The 2N-order eigensolve uses numpy's LAPACK zgeev (results do not depend on eigenvector order or
phase).  Block-diagonal 2 x 2 operators are stored as a (4, N) array [a11, a12, a21, a22].
"""
from __future__ import annotations

import math

import numpy as np

from .flat import decode2d

PI = 3.141592653589793238462643383279502884


class NumericalFailure(Exception):
    pass


class RayleighFailure(Exception):
    pass


# ---- operators ------------------------------------------------------------------------------------
def fourier_1d(edges, values, period, hmax):
    e = np.asarray(edges, dtype=float)
    v = np.asarray(values, dtype=complex)
    c = np.zeros(2 * hmax + 1, dtype=complex)
    c[hmax] = np.sum(v * (e[1:] - e[:-1])) / period
    for h in range(-hmax, hmax + 1):
        if h == 0:
            continue
        arg = (-2.0 * PI * (h * e)) / period
        ph = np.cos(arg) + 1j * np.sin(arg)
        s = np.sum(v * (ph[1:] - ph[:-1]))
        w = -2.0 * PI * h
        c[h + hmax] = complex(s.imag / w, -s.real / w)
    return c


def toeplitz_1d(c, M):
    n = 2 * M + 1
    idx = np.arange(n)[:, None] - np.arange(n)[None, :] + 2 * M
    return c[idx]


def seg_fourier(edges, hmax, period):
    e = np.asarray(edges, dtype=float)
    K = e.size - 1
    out = np.zeros((K, 2 * hmax + 1), dtype=complex)
    out[:, hmax] = (e[1:] - e[:-1]) / period
    for h in range(-hmax, hmax + 1):
        if h == 0:
            continue
        arg = (-2.0 * PI * (h * e)) / period
        ph = np.cos(arg) + 1j * np.sin(arg)
        d = ph[1:] - ph[:-1]
        w = -2.0 * PI * h
        out[:, h + hmax] = d.imag / w - 1j * (d.real / w)
    return out


def floor_mod(a, b):
    m = math.fmod(a, b)
    if m != 0.0:
        if (b < 0.0) != (m < 0.0):
            m += b
    else:
        m = math.copysign(0.0, b)
    return m


def eps_at(prof, x, Lx):
    xe = prof.edges
    xx = xe[0] + floor_mod(x - xe[0], Lx)
    k = int(np.searchsorted(np.asarray(xe), xx, side="right")) - 1
    k = min(max(k, 0), len(prof.values) - 1)
    n = prof.values[k]
    return n * n


def merge_x_cuts(rows, Lx):
    cuts = sorted(set([floor_mod(e, Lx) for p in rows for e in p.edges] + [0.0, Lx]))
    cuts = sorted(set(float(np.rint(c * 1e12) / 1e12) for c in cuts))
    cuts = [c for c in cuts if -1e-12 <= c <= Lx + 1e-12]
    if not cuts or cuts[0] > 1e-12:
        cuts.insert(0, 0.0)
    if cuts[-1] < Lx - 1e-12:
        cuts.append(Lx)
    return cuts


def assemble_xfirst(blocks, M1, M2):
    n1, n2 = 2 * M1 + 1, 2 * M2 + 1
    N = n1 * n2
    out = np.zeros((N, N), dtype=complex)
    for col in range(N):
        mp, np_ = divmod(col, n2)
        for nn in range(n2):
            src = blocks[nn - np_ + 2 * M2][:, mp]
            out[np.arange(n1) * n2 + nn, col] = src
    return out


def assemble_yfirst(blocks, M1, M2):
    n1, n2 = 2 * M1 + 1, 2 * M2 + 1
    N = n1 * n2
    out = np.zeros((N, N), dtype=complex)
    for col in range(N):
        mp, np_ = divmod(col, n2)
        for mm in range(n1):
            src = blocks[mm - mp + 2 * M1][:, np_]
            out[mm * n2 + np.arange(n2), col] = src
    return out


def _inv(a):
    try:
        return np.linalg.solve(a, np.eye(a.shape[0], dtype=complex))
    except np.linalg.LinAlgError as exc:
        raise NumericalFailure() from exc


def layer_operators(layer, Lx, Ly, M1, M2, laurent=False):
    n1, n2 = 2 * M1 + 1, 2 * M2 + 1
    K = len(layer.y_edges) - 1
    rows = layer.rows
    wseg = seg_fourier(layer.y_edges, 2 * M2, Ly)
    lau = [toeplitz_1d(fourier_1d(p.edges, [v * v for v in p.values], Lx, 2 * M1), M1) for p in rows]
    eps_h = [sum(wseg[k, h] * lau[k] for k in range(K)) for h in range(4 * M2 + 1)]
    eps = assemble_xfirst(eps_h, M1, M2)
    if laurent:
        return eps, eps.copy(), eps.copy()
    invx = [_inv(toeplitz_1d(fourier_1d(p.edges, [1.0 / (v * v) for v in p.values], Lx, 2 * M1), M1)) for p in rows]
    cx_h = [sum(wseg[k, h] * invx[k] for k in range(K)) for h in range(4 * M2 + 1)]
    cx = assemble_xfirst(cx_h, M1, M2)
    xcuts = merge_x_cuts(rows, Lx)
    Kx = len(xcuts) - 1
    wsegx = seg_fourier(xcuts, 2 * M1, Lx)
    invy = []
    for j in range(Kx):
        xm = 0.5 * (xcuts[j] + xcuts[j + 1])
        inv_col = [1.0 / eps_at(rows[k], xm, Lx) for k in range(K)]
        invy.append(_inv(toeplitz_1d(fourier_1d(layer.y_edges, inv_col, Ly, 2 * M2), M2)))
    cy_h = [sum(wsegx[j, h] * invy[j] for j in range(Kx)) for h in range(4 * M1 + 1)]
    cy = assemble_yfirst(cy_h, M1, M2)
    return eps, cx, cy


# ---- modes ------------------------------------------------------------------------------------------
def branch(q):
    q = np.where(q.real < 0.0, -q, q)
    return np.where((q.real < 1e-10) & (q.imag > 0.0), -q, q)


def kz_over_k0(n, Kx, Ky):
    kz = np.sqrt(n * n - Kx * Kx - Ky * Ky + 0j)
    return np.where(kz.imag < 0.0, -kz, kz)


class Modes:
    def __init__(self):
        self.homog = False
        self.q = None
        self.Vb = None
        self.W = self.V = None


def homogeneous_modes(n, Kx, Ky, kz):
    m = Modes()
    m.homog = True
    q1 = branch(kz.imag - 1j * kz.real)
    m.q = np.concatenate([q1, q1])
    n2 = n * n
    m.Vb = np.stack([(Kx * Ky) / q1, (n2 - Kx * Kx) / q1, (Ky * Ky - n2) / q1, (-Ky * Kx) / q1])
    return m


def structured_modes(eps, cx, cy, Kx, Ky):
    N = Kx.size
    Ei = _inv(eps)
    I = np.eye(N)
    P = np.block([[Kx[:, None] * Ei * Ky[None, :], I - Kx[:, None] * Ei * Kx[None, :]],
                  [Ky[:, None] * Ei * Ky[None, :] - I, -(Ky[:, None] * Ei * Kx[None, :])]])
    q12 = cy - np.diag(Kx * Kx)
    q21 = np.diag(Ky * Ky) - cx
    Q = np.block([[np.diag(Kx * Ky), q12], [q21, np.diag(-(Ky * Kx))]])
    H = P @ Q
    lam, W = np.linalg.eig(H)
    W = W / np.linalg.norm(W, axis=0)[None, :]
    m = Modes()
    m.q = branch(np.sqrt(lam))
    m.W = W
    m.V = (Q @ W) / m.q[None, :]
    if not (np.all(np.isfinite(m.W)) and np.all(np.isfinite(m.V))):
        raise NumericalFailure()
    return m


# ---- 2 x 2 block-diagonal algebra --------------------------------------------------------------------
def b2_mul(A, B):
    return np.stack([A[0] * B[0] + A[1] * B[2], A[0] * B[1] + A[1] * B[3],
                     A[2] * B[0] + A[3] * B[2], A[2] * B[1] + A[3] * B[3]])


def b2_inv(A):
    det = A[0] * A[3] - A[1] * A[2]
    return np.stack([A[3] / det, -A[1] / det, -A[2] / det, A[0] / det])


def b2_eye(N):
    return np.stack([np.ones(N, complex), np.zeros(N, complex), np.zeros(N, complex), np.ones(N, complex)])


def b2_dense(A):
    N = A.shape[1]
    m = np.zeros((2 * N, 2 * N), complex)
    i = np.arange(N)
    m[i, i] = A[0]; m[i, N + i] = A[1]; m[N + i, i] = A[2]; m[N + i, N + i] = A[3]
    return m


def b2_apply(A, v):
    N = A.shape[1]
    return np.concatenate([A[0] * v[:N] + A[1] * v[N:], A[2] * v[:N] + A[3] * v[N:]])


def b2_rows(A, M):
    N = A.shape[1]
    return np.vstack([A[0][:, None] * M[:N] + A[1][:, None] * M[N:], A[2][:, None] * M[:N] + A[3][:, None] * M[N:]])


def b2_rmul(M, B):
    N = B.shape[1]
    return np.hstack([M[:, :N] * B[0][None, :] + M[:, N:] * B[2][None, :],
                      M[:, :N] * B[1][None, :] + M[:, N:] * B[3][None, :]])


def _solve(a, b):
    try:
        return np.linalg.solve(a, b)
    except np.linalg.LinAlgError as exc:
        raise NumericalFailure() from exc


def interface_bd(up, low):
    N = up.Vb.shape[1]
    nn = b2_mul(b2_inv(low.Vb), up.Vb)
    mm = b2_eye(N)
    tp = 0.5 * (mm + nn); tm = 0.5 * (mm - nn)
    tpi = b2_inv(tp)
    return dict(tdd=tp - b2_mul(b2_mul(tm, tpi), tm), rdu=b2_mul(tm, tpi), rud=-b2_mul(tpi, tm), tuu=tpi)


def interface_matching(up, low):
    if low.homog:
        return up.W.copy(), b2_rows(b2_inv(low.Vb), up.V)
    if up.homog:
        return _inv(low.W), _solve(low.V, b2_dense(up.Vb))
    return _solve(low.W, up.W), _solve(low.V, up.V)


def solve(stack_flat, wavelength, theta_deg, phi_deg, M1, M2, jones, laurent=False):
    """One crossed-grating solve; returns dict(status, DEr, DEt, Rs, Rp, Ts, Tp, Ex_r, Ey_r, Ex_t, Ey_t)."""
    st = decode2d(stack_flat)
    n1, n2 = 2 * M1 + 1, 2 * M2 + 1
    N = n1 * n2
    n = 2 * N
    L = len(st.layers)
    nI, nII = st.nI, st.nII
    k0 = 2.0 * PI / wavelength
    th, ph = theta_deg * (PI / 180.0), phi_deg * (PI / 180.0)
    ct, s_t, cp, sp = math.cos(th), math.sin(th), math.cos(ph), math.sin(ph)
    mm_, nn_ = np.divmod(np.arange(N), n2)
    m_idx = mm_ - M1
    n_idx = nn_ - M2
    Kx = nI * s_t * cp - m_idx * wavelength / st.Lx
    Ky = nI * s_t * sp - n_idx * wavelength / st.Ly
    kzI = kz_over_k0(nI, Kx, Ky)
    kzII = kz_over_k0(nII, Kx, Ky)
    nan = np.full(N, np.nan)
    cnan = np.full(N, complex(np.nan, np.nan))
    fail = dict(DEr=nan, DEt=nan, Rs=cnan, Rp=cnan, Ts=cnan, Tp=cnan, Ex_r=cnan, Ey_r=cnan, Ex_t=cnan, Ey_t=cnan)
    try:
        if np.any(kzI == 0) or np.any(kzII == 0):
            raise RayleighFailure()
        for layer in st.layers:
            if layer.homogeneous and np.any(kz_over_k0(layer.n, Kx, Ky) == 0):
                raise RayleighFailure()
        modes = [None] * (L + 2)
        modes[0] = homogeneous_modes(nI, Kx, Ky, kzI)
        modes[-1] = homogeneous_modes(nII, Kx, Ky, kzII)
        for j, layer in enumerate(st.layers):
            if layer.homogeneous:
                modes[j + 1] = homogeneous_modes(layer.n, Kx, Ky, kz_over_k0(layer.n, Kx, Ky))
            else:
                eps, cx, cy = layer_operators(layer, st.Lx, st.Ly, M1, M2, laurent)
                modes[j + 1] = structured_modes(eps, cx, cy, Kx, Ky)
        inc = np.zeros(n, complex)
        i0 = M1 * n2 + M2
        Es, Ep = jones
        inc[i0] = -Es * sp + Ep * ct * cp
        inc[N + i0] = Es * cp + Ep * ct * sp
        # bottom-up: r = ("z",) | ("bd", B) | ("dense", R); steps
        r = ("z", None)
        steps = [None] * (2 * L + 1)
        for b in range(2 * L, 0, -1):
            j = b // 2
            if b % 2 == 0:
                up, low = modes[j], modes[j + 1]
                if up.homog and low.homog:
                    if r[0] != "dense":
                        blk = interface_bd(up, low)
                        rr = np.zeros((4, N), complex) if r[0] == "z" else r[1]
                        I_minus = b2_eye(N) - b2_mul(blk["rdu"], rr)
                        gt = b2_mul(b2_inv(I_minus), blk["tdd"])
                        out = blk["rud"] if r[0] == "z" else blk["rud"] + b2_mul(b2_mul(blk["tuu"], rr), gt)
                        steps[b] = ("bd", gt)
                        r = ("bd", out)
                    else:
                        nnb = b2_mul(b2_inv(low.Vb), up.Vb)
                        one = b2_eye(N)
                        np1, nm1 = nnb + one, nnb - one
                        ap = b2_rmul(r[1], nm1) + b2_dense(np1)
                        am = b2_rmul(r[1], np1) + b2_dense(nm1)
                        rp = _solve(ap, am)
                        steps[b] = ("iface_bd", -0.5 * nm1, nnb, rp)
                        r = ("dense", rp)
                else:
                    mm, nn = interface_matching(up, low)
                    if r[0] == "z":
                        pn, qm = nn, mm
                    elif r[0] == "bd":
                        pn = b2_rows(b2_eye(N) + r[1], nn)
                        qm = b2_rows(b2_eye(N) - r[1], mm)
                    else:
                        pn = r[1] @ nn + nn
                        qm = mm - r[1] @ mm
                    rp = _solve(pn + qm, pn - qm)
                    steps[b] = ("iface", 0.5 * (mm - nn), nn, rp)
                    r = ("dense", rp)
            else:
                q = modes[j + 1].q
                d = st.layers[j].d
                x = np.exp(-k0 * q * d)
                tdd = np.stack([x[:N], np.zeros(N, complex), np.zeros(N, complex), x[N:]])
                if r[0] != "dense":
                    rr = np.zeros((4, N), complex) if r[0] == "z" else r[1]
                    out = np.zeros((4, N), complex) if r[0] == "z" else b2_mul(b2_mul(tdd, rr), tdd)
                    steps[b] = ("bd", tdd)
                    r = ("bd", out)
                else:
                    steps[b] = ("bd", tdd)
                    r = ("dense", b2_rmul(b2_rows(tdd, r[1]), tdd))
        # top block (direct form)
        up, low = modes[0], modes[1]
        if up.homog and low.homog:
            blk = interface_bd(up, low)
            if r[0] != "dense":
                rr = np.zeros((4, N), complex) if r[0] == "z" else r[1]
                g = b2_inv(b2_eye(N) - b2_mul(blk["rdu"], rr))
                down = b2_apply(b2_mul(g, blk["tdd"]), inc)
                back = b2_apply(blk["tuu"], b2_apply(rr, down))
            else:
                mat = np.eye(n) - b2_rows(blk["rdu"], r[1])
                down = _solve(mat, b2_apply(blk["tdd"], inc))
                back = b2_apply(blk["tuu"], r[1] @ down)
            R = b2_apply(blk["rud"], inc) + back
        else:
            if r[0] == "z":
                wp, vq = low.W, low.V
            elif r[0] == "bd":
                wp = b2_rmul(low.W, b2_eye(N) + r[1]); vq = b2_rmul(low.V, b2_eye(N) - r[1])
            else:
                wp = low.W @ r[1] + low.W; vq = low.V - low.V @ r[1]
            K = b2_rows(b2_inv(up.Vb), vq) + wp
            down = _solve(K, 2.0 * inc)
            R = wp @ down - inc
        for b in range(1, 2 * L + 1):
            s = steps[b]
            if s[0] == "bd":
                down = b2_apply(s[1], down)
            elif s[0] == "iface_bd":
                u = s[3] @ down + down
                down = b2_apply(s[1], u) + b2_apply(s[2], down)
            else:
                u = s[3] @ down + down
                down = s[1] @ u + s[2] @ down
        T = down
        if not (np.all(np.isfinite(R)) and np.all(np.isfinite(T))):
            raise NumericalFailure()

        def efficiency(F, kz, sign):
            Ex, Ey = F[:N].copy(), F[N:].copy()
            Ez = np.zeros(N, complex); de = np.zeros(N)
            for i in range(N):
                if not (kz[i].real > 1e-12):
                    continue
                Ez[i] = -sign * (Kx[i] * Ex[i] + Ky[i] * Ey[i]) / kz[i]
                de[i] = kz[i].real * (abs(Ex[i]) ** 2 + abs(Ey[i]) ** 2 + abs(Ez[i]) ** 2) / (nI * ct)
            return Ex, Ey, de

        def spol(Ex, Ey, kz, nmed, sign):
            s_ = np.zeros(N, complex); p_ = np.zeros(N, complex)
            for i in range(N):
                kt = math.hypot(Kx[i], Ky[i])
                cth = kz[i] / nmed if abs(kz[i]) > 1e-12 else 1.0
                if kt > 1e-12:
                    s_[i] = (-Ky[i] * Ex[i] + Kx[i] * Ey[i]) / kt
                    p_[i] = sign * (Kx[i] * Ex[i] + Ky[i] * Ey[i]) / (kt * cth)
                else:
                    s_[i] = -sp * Ex[i] + cp * Ey[i]
                    p_[i] = sign * (cp * Ex[i] + sp * Ey[i]) / cth
            return s_, p_

        Exr, Eyr, DEr = efficiency(R, kzI, -1.0)
        Ext, Eyt, DEt = efficiency(T, kzII, 1.0)
        Rs, Rp = spol(Exr, Eyr, kzI, nI, -1.0)
        Ts, Tp = spol(Ext, Eyt, kzII, nII, 1.0)
        return dict(status=0, DEr=DEr, DEt=DEt, Rs=Rs, Rp=Rp, Ts=Ts, Tp=Tp, Ex_r=Exr, Ey_r=Eyr, Ex_t=Ext, Ey_t=Eyt)
    except NumericalFailure:
        return dict(status=1, **fail)
    except RayleighFailure:
        return dict(status=2, **fail)
