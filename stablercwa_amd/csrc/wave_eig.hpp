// Batched complex nonsymmetric eigensolver with one 64-lane wave per matrix (small orders).
//
// Same algorithm as wg_eig (the reference's eig_qr, include/stablercwa/eig_qr.hpp:1008-1124: CORTH
// Householder reduction, accumulation of Z, real sub-diagonal, COMQR2 single-shift QR with the same
// shifts, deflation test and iteration limit, back-substitution with the reference's denominator floor
// and overflow control, V = Z X with unit columns).  What changes is the mapping onto the chip:
//
//   * A workgroup of 256 threads holds four independent matrices, one per wave, and no step needs a
//     workgroup barrier.  The QR sweep is a chain of dependent rotations; with the chain inside one
//     wave the rotation is broadcast by readlane instead of an LDS round trip plus barrier, and four
//     times as many matrices are resident per CU to hide the chain's latency.
//   * Householder and accumulation steps: one lane per column (a serial dot product down the column,
//     no cross-lane reduction), one lane per row for the right-hand update.
//   * QR sweep, left rotations: lane c owns columns l + c + 64 q; each column's next rows are loaded
//     WAVE_RING rotations ahead into registers.  Right rotations: one lane per row of H and Z.
//   * Back-substitution: one lane per eigenvector; V = Z X on MFMA tiles computed by the wave alone.
#pragma once

#include "eig.hpp"

// Orders up to which the eigen kernels use one wave per matrix.  Measured on MI355X for batches of
// thousands of matrices: at n = 61 the wave form is 3x faster than a workgroup per matrix (the
// matrices of four waves stay in cache and no rotation waits at a workgroup barrier); at n = 121 and
// above the workgroup form wins (the wave form's larger in-flight working set streams from HBM).
constexpr int WAVE_EIG_MAX_N = SRCWA_WAVE_EIG_MAX_N;
constexpr int WAVE_RING = 4;

// Columns per lane for matrices of order up to n (wave kernels are instantiated for 1, 2 and 4).
inline int wave_columns_per_lane(int n) { return n <= 64 ? 1 : n <= 128 ? 2 : 4; }


// Eigen-decomposition of the n x n matrix A by the calling wave (all 64 lanes must call it).
//   QC   : columns per lane in the QR sweep (n <= 64 QC)
//   rot  : 3 (n + 1) doubles of LDS owned by this wave (the sweep's rotations)
// Outputs as wg_eig: A = unit-norm right eigenvectors, w = eigenvalues, info (0 or k > 0 when
// eigenvalue k did not converge); vectors == false leaves the Schur form T in A and Z.
template <int QC>
DEV int wave_eig(int n, cd* __restrict__ A, cd* __restrict__ w, cd* __restrict__ Z, cd* __restrict__ X,
                 EigScratch s, double* rot, int max_iter, double eps, bool vectors = true) {
    if (n > 64 * QC) return n + 1;
    const int lane = lane_id();
    cd* H = A;
    cd* ort = s.ort;
    // orders are at most 256, so every element offset fits an int (32-bit offsets from a uniform base)
    auto h = [&](int i, int j) -> cd& { return H[i + j * n]; };
    double* rot_c = rot;
    double* rot_ci = rot + (n + 1);
    double* rot_s = rot + 2 * (n + 1);

    // ---- 1. Householder reduction to Hessenberg form (CORTH)
    for (int m = 1; m < n - 1; ++m) {
        cd* x = H + (m - 1) * n;
        double part = 0.0;
        for (int i = m + lane; i < n; i += 64) part += abs1(x[i]);
        const double scale = wave_sum(part);
        if (scale == 0.0) {
            if (lane == 0) ort[m] = cd{0.0, 0.0};
            wave_sync();
            continue;
        }
        double hp = 0.0;
        for (int i = m + lane; i < n; i += 64) {
            const cd u = cdiv(x[i], scale);
            ort[i] = u;
            hp += norm2(u);
        }
        double hh = wave_sum(hp);
        double g = sqrt(hh);
        wave_sync();
        const cd om = ort[m];
        const double f = cabs(om);
        if (f == 0.0) {
            if (lane == 0) {
                ort[m] = cd{g, 0.0};
                x[m] = cd{scale, 0.0};
            }
        } else {
            hh += f * g;
            g /= f;
            if (lane == 0) ort[m] = om * (1.0 + g);
        }
        wave_sync();
        // left: one lane per column j in [m, n)
        for (int j = m + lane; j < n; j += 64) {
            cd* aj = H + j * n;
            double sr = 0.0, si = 0.0;
#pragma unroll 2
            for (int i = m; i < n; ++i) {
                const cd u = ort[i], a = aj[i];
                sr += u.x * a.x + u.y * a.y;
                si += u.x * a.y - u.y * a.x;
            }
            const cd t = cd{-sr / hh, -si / hh};
#pragma unroll 2
            for (int i = m; i < n; ++i) aj[i] = aj[i] + t * ort[i];
        }
        wave_sync();
        // right: one lane per row
        for (int i = lane; i < n; i += 64) {
            cd acc{0.0, 0.0};
#pragma unroll 2
            for (int j = n - 1; j >= m; --j) acc = acc + ort[j] * h(i, j);
            acc = cdiv(acc, hh);
#pragma unroll 2
            for (int j = m; j < n; ++j) h(i, j) = h(i, j) - conj(ort[j]) * acc;
        }
        wave_sync();
        if (lane == 0) {
            ort[m] = ort[m] * scale;
            x[m] = x[m] * (-g);
        }
        wave_sync();
    }

    // ---- 2. Z = P_1 ... P_{n-2}
    for (int e = lane; e < n * n; e += 64) Z[e] = cd{(e % n) == (e / n) ? 1.0 : 0.0, 0.0};
    cd* v = s.tmp;
    wave_sync();
    for (int i = n - 2; i >= 1; --i) {
        const cd oi = ort[i];
        const cd sub0 = h(i, i - 1);
        if (is_zero(oi) || is_zero(sub0)) continue;
        const int e = binary_exponent(abs1(oi));
        const cd sub = scale2(sub0, -e);
        const cd v0 = scale2(oi, -e);
        const double minus_h = sub.x * v0.x + sub.y * v0.y;
        const int len = n - i;
        for (int k = lane; k < len; k += 64) v[k] = k == 0 ? v0 : scale2(h(i + k, i - 1), -e);
        wave_sync();
        for (int j = i + lane; j < n; j += 64) {
            cd* zj = Z + j * n + i;
            double sr = 0.0, si = 0.0;
#pragma unroll 2
            for (int k = 0; k < len; ++k) {
                const cd u = v[k], a = zj[k];
                sr += u.x * a.x + u.y * a.y;
                si += u.x * a.y - u.y * a.x;
            }
            const cd t = cd{sr / minus_h, si / minus_h};
#pragma unroll 2
            for (int k = 0; k < len; ++k) zj[k] = zj[k] + t * v[k];
        }
        wave_sync();
    }

    // ---- 3. clear below the sub-diagonal; make the sub-diagonal real (D^H H D, Z D)
    for (int e = lane; e < n * n; e += 64) {
        const int i = static_cast<int>(e % n), j = static_cast<int>(e / n);
        if (i >= j + 2) h(i, j) = cd{0.0, 0.0};
    }
    wave_sync();
    cd* y = s.tmp;
    if (lane == 0) {
        cd prev{1.0, 0.0};
        bool prev_scaled = false;
        y[0] = cd{1.0, 0.0};
        for (int i = 1; i < n; ++i) {
            cd sub = h(i, i - 1);
            if (prev_scaled) sub = sub * prev;
            if (sub.y == 0.0) {
                h(i, i - 1) = sub;
                y[i] = cd{1.0, 0.0};
                prev_scaled = false;
            } else {
                double norm = 0.0;
                const cd yy = unit_phase(sub, norm);
                h(i, i - 1) = cd{norm, 0.0};
                y[i] = yy;
                prev = yy;
                prev_scaled = true;
            }
        }
    }
    wave_sync();
    for (int e = lane; e < n * n; e += 64) {
        const int i = static_cast<int>(e % n), j = static_cast<int>(e / n);
        if (i <= j) {
            cd val = h(i, j);
            if (i >= 1) val = conj(y[i]) * val;
            val = val * y[j];
            h(i, j) = val;
        }
        Z[e] = Z[e] * y[j];
    }
    wave_sync();

    // ---- 4. shifted QR iterations (COMQR2)
    cd t{0.0, 0.0};
    int en = n - 1, its = 0, info = 0;
    while (en >= 0) {
        int lc = 0;
        for (int l = 1 + lane; l <= en; l += 64)
            if (abs1(h(l, l - 1)) <= eps * (abs1(h(l - 1, l - 1)) + abs1(h(l, l)))) lc = l;
        const int l = wave_max_int(lc);
        if (l == en) {
            if (lane == 0) {
                const cd val = h(en, en) + t;
                w[en] = val;
                h(en, en) = val;
            }
            wave_sync();
            --en;
            its = 0;
            continue;
        }
        if (its == max_iter) {
            info = en + 1;
            break;
        }
        cd sh;
        if (its == 10 || its == 20) {
            sh = cd{fabs(h(en, en - 1).x) + (en >= 2 ? fabs(h(en - 1, en - 2).x) : 0.0), 0.0};
        } else {
            const cd a = h(en - 1, en - 1), b = h(en - 1, en), d = h(en, en);
            const double c = h(en, en - 1).x;
            const cd yv = 0.5 * (a - d);
            const int e = binary_exponent(fmax(fmax(abs1(b), fabs(c)), abs1(yv)));
            const cd xv = scale2(b, -e) * ldexp(c, -e);
            if (is_zero(xv)) {
                sh = d;
            } else {
                const cd ys = scale2(yv, -e);
                cd zz = csqrt(ys * ys + xv);
                if (ys.x * zz.x + ys.y * zz.y < 0.0) zz = -zz;
                sh = d - scale2(cdiv(xv, ys + zz), e);
            }
        }
        wave_sync();
        for (int i = lane; i <= en; i += 64) h(i, i) = h(i, i) - sh;
        t = t + sh;
        ++its;
        wave_sync();

        // (1) left rotations
        wave_left_sweep<QC, WAVE_RING>(n, l, en, H, rot_c, rot_ci, rot_s);
        wave_sync();

        // (2) phase that makes r(en, en) real
        const cd hee = h(en, en);
        const bool has_phase = hee.y != 0.0;
        cd phase{1.0, 0.0};
        if (has_phase) {
            double nrm = 0.0;
            phase = unit_phase(hee, nrm);
            wave_sync();
            if (lane == 0) h(en, en) = cd{nrm, 0.0};
            for (int j = en + 1 + lane; j < n; j += 64) h(en, j) = conj(phase) * h(en, j);
        }
        wave_sync();

        // (3) right rotations: rows of H (0..en) and Z (0..n-1), one lane per row
        constexpr int RB = 4;
        for (int r = lane; r < en + 1 + n; r += 64) {
            const bool zrow = r > en;
            const int row = zrow ? r - en - 1 : r;
            cd* base = zrow ? Z : H;
            const int j0 = zrow ? l + 1 : max(l + 1, row);
            cd a = base[row + (j0 - 1) * n];
            for (int jb = j0; jb <= en; jb += RB) {
                cd bv[RB];
#pragma unroll
                for (int u = 0; u < RB; ++u)
                    if (jb + u <= en) bv[u] = base[row + (jb + u) * n];
#pragma unroll
                for (int u = 0; u < RB; ++u) {
                    const int j = jb + u;
                    if (j <= en) {
                        const double cr = rot_c[j], ci = rot_ci[j], sn = rot_s[j];
                        const cd b = bv[u];
                        base[row + (j - 1) * n] =
                            cd{cr * a.x - ci * a.y + sn * b.x, cr * a.y + ci * a.x + sn * b.y};
                        a = cd{cr * b.x + ci * b.y - sn * a.x, cr * b.y - ci * b.x - sn * a.y};
                    }
                }
            }
            if (has_phase) a = a * phase;
            base[row + en * n] = a;
        }
        wave_sync();
    }
    if (info != 0) {
        for (int i = lane; i <= en; i += 64) w[i] = h(i, i) + t;
        wave_sync();
        return info;
    }

    // ---- 5. eigenvectors of T, V = Z X, normalise
    for (int j = lane; j + 1 < n; j += 64) h(j + 1, j) = cd{0.0, 0.0};
    wave_sync();
    if (!vectors) return 0;
    double nm = 0.0;
    for (int e = lane; e < n * n; e += 64) {
        const int i = static_cast<int>(e % n), j = static_cast<int>(e / n);
        if (i <= j) nm = fmax(nm, abs1(H[e]));
    }
    const double tnorm = wave_max(nm);
    const double rel = fmax(eps, kEps);
    double tiny = rel * tnorm;
    if (tiny == 0.0) tiny = 2.2250738585072014e-308;
    constexpr double kQuotientMax = 1e300;
    const double kShrink = ldexp(1.0, -512);
    // X row-major: X[k*n + j] = component k of eigenvector j
    for (int e = lane; e < n * n; e += 64) {
        const int k = static_cast<int>(e / n), j = static_cast<int>(e % n);
        X[e] = k < j ? h(k, j) : (k == j ? cd{1.0, 0.0} : cd{0.0, 0.0});
    }
    wave_sync();
    for (int j = lane; j < n; j += 64) {
        const cd wj = w[j];
        for (int i = j - 1; i >= 0; --i) {
            const cd wi = w[i];
            cd yv = wj - wi;
            if (is_zero(yv)) {
                yv = cd{tiny, 0.0};
            } else {
                const double fl = rel * fmax(abs1(wj), abs1(wi));
                if (abs1(yv) < fl) yv = cd{fl, 0.0};
            }
            for (int pass = 0; pass < 3 && abs1(X[i * n + j]) > abs1(yv) * kQuotientMax; ++pass)
                for (int k = 0; k <= j; ++k) X[k * n + j] = X[k * n + j] * kShrink;
            const cd xi = cdiv(X[i * n + j], yv);
            X[i * n + j] = xi;
            const double tr = abs1(xi);
            cd xs = xi;
            if (tr != 0.0 && tr + 1.0 / tr <= tr) {
                for (int k = 0; k <= j; ++k) X[k * n + j] = cdiv(X[k * n + j], tr);
                xs = X[i * n + j];
            }
            const cd* ti = H + i * n;
            for (int k = 0; k < i; ++k) X[k * n + j] = X[k * n + j] + xs * ti[k];
        }
    }
    wave_sync();
    // V = Z X into A (T is no longer needed)
    wg_gemm<GEMM_SET, true, true>(n, n, n, Z, n, X, n, A, n);
    wave_sync();
    for (int j = lane; j < n; j += 64) {
        cd* vj = A + j * n;
        double big = 0.0;
        for (int i = 0; i < n; ++i) big = fmax(big, fmax(fabs(vj[i].x), fabs(vj[i].y)));
        if (big == 0.0 || !isfinite(big)) continue;
        double ssq = 0.0;
        for (int i = 0; i < n; ++i) ssq += norm2(cdiv(vj[i], big));
        const double inv = 1.0 / sqrt(ssq);
        for (int i = 0; i < n; ++i) vj[i] = cdiv(vj[i], big) * inv;
    }
    wave_sync();
    return 0;
}
