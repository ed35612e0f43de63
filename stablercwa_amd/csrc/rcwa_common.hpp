// Pieces of the planar / conical RCWA solve shared by the kernels: flat stack parsing, analytic
// Fourier coefficients and Toeplitz matrices, homogeneous-medium modes, and the S-matrix
// recursion with the reference's zero / diagonal / dense operator algebra.
#pragma once

#include "batch.hpp"
#include "tuned.hpp"
#include "wg_la.hpp"
#if SRCWA_MIXED_PRECISION
#include "wg_la32.hpp"
#endif

constexpr int SMAT_MAX_LAYERS = 128;
constexpr int SMAT_MAX_BLOCKS = 2 * SMAT_MAX_LAYERS + 1;

// ------------------------------------------------------------------------------------------------
// Flat planar stack record (see include/srcwa.h): period, nI, nII, L, then per layer
// d, kind, uniform: n | grating: K, edges[K+1], n[K], method, samples.
struct LayerView {
    double d;
    bool homogeneous;
    cd n;
    int K;
    const double* edges;
    const double* vals;     // K complex, interleaved
    int method;             // gratings: 0 analytic, 1 FFT (FourierMethod)
    int samples;            // gratings: fft_samples (0 automatic)
};

struct LayerCursor {
    const double* p;
    int64_t k;
    DEV explicit LayerCursor(const double* st) : p(st), k(6) {}
    DEV LayerView next() {
        LayerView v;
        v.d = p[k];
        const int kind = (int)p[k + 1];
        k += 2;
        if (kind == 0) {
            v.homogeneous = true;
            v.n = cd{p[k], p[k + 1]};
            v.K = 1;
            v.edges = nullptr;
            v.vals = nullptr;
            v.method = 0;
            v.samples = 0;
            k += 2;
        } else {
            v.homogeneous = false;
            v.n = cd{1.0, 0.0};
            v.K = (int)p[k];
            v.edges = p + k + 1;
            v.vals = p + k + 2 + v.K;
            v.method = (int)v.vals[2 * v.K];
            v.samples = (int)v.vals[2 * v.K + 1];
            k += 1 + (v.K + 1) + 2 * v.K + 2;
        }
        return v;
    }
};

// c_h, h = -hmax..hmax, of eps = n^2 (or 1/eps when inverse), stored at out[h + hmax]
// (rcwa1d.hpp:366-388; a uniform layer is one segment on [0, period]).
DEV void fourier_coefficients(const LayerView& lv, double period, int hmax, bool inverse, cd* out) {
    for (int hh = threadIdx.x; hh <= 2 * hmax; hh += NT) {
        const int h = hh - hmax;
        cd sum{0.0, 0.0};
        const int K = lv.K;
        if (h == 0) {
            for (int j = 0; j < K; ++j) {
                const cd nj = lv.homogeneous ? lv.n : cd{lv.vals[2 * j], lv.vals[2 * j + 1]};
                cd v = nj * nj;
                if (inverse) v = cdiv(cd{1.0, 0.0}, v);
                const double e0 = lv.homogeneous ? 0.0 : lv.edges[j];
                const double e1 = lv.homogeneous ? period : lv.edges[j + 1];
                sum = sum + cdiv(v * (e1 - e0), period);
            }
        } else {
            const double w = -2.0 * kPi * h;
            double prev_s, prev_c;
            {
                const double e0 = lv.homogeneous ? 0.0 : lv.edges[0];
                sincos(-2.0 * kPi * h * e0 / period, &prev_s, &prev_c);
            }
            for (int j = 0; j < K; ++j) {
                const cd nj = lv.homogeneous ? lv.n : cd{lv.vals[2 * j], lv.vals[2 * j + 1]};
                cd v = nj * nj;
                if (inverse) v = cdiv(cd{1.0, 0.0}, v);
                const double e1 = lv.homogeneous ? period : lv.edges[j + 1];
                double s1, c1;
                sincos(-2.0 * kPi * h * e1 / period, &s1, &c1);
                const cd t = v * cd{c1 - prev_c, s1 - prev_s};
                sum = sum + cd{t.y / w, -t.x / w};
                prev_s = s1;
                prev_c = c1;
            }
        }
        out[hh] = sum;
    }
    __syncthreads();
}

// eps = n^2 (or 1/eps) of segment j of a grating
DEV cd segment_value(const LayerView& lv, int j, bool inverse) {
    const cd nj{lv.vals[2 * j], lv.vals[2 * j + 1]};
    const cd v = nj * nj;
    return inverse ? cdiv(cd{1.0, 0.0}, v) : v;
}

// c_h, h = -hmax..hmax, of a grating's eps (or independently sampled 1/eps) from nfft samples at the cell
// midpoints u_j = (j + 1/2) / nfft * period: a forward DFT scaled by 1/nfft, then the phase of the grating's
// origin and of the half cell, exp(-i 2 pi h (origin / period + 1 / (2 nfft))) (fourier.hpp:46-83).  A
// power-of-two nfft takes a radix-2 Stockham FFT over buf[0, 2 nfft); any other nfft a direct DFT of the
// 2 hmax + 1 bins that are needed.  Returns 1 when a sampled value is not finite (the reference's numerical
// failure, rcwa1d.hpp:393-398).
DEV int fft_coefficients(const LayerView& lv, double period, int hmax, bool inverse, int nfft, cd* buf, cd* out) {
    const int K = lv.K;
    const double e0 = lv.edges[0];
    int bad = 0;
    for (int k = threadIdx.x; k < K; k += NT) bad |= !finite(segment_value(lv, k, inverse));
    if (block_max_int(bad)) return 1;
    // the segment of u: the number of interior edges k with u >= edges[k] - edges[0] (a binary search
    // over the same comparisons as the reference's walk)
    auto sample = [&](int j) -> cd {
        const double u = (static_cast<double>(j) + 0.5) / nfft * period;
        int lo = 0, hi = K - 1;            // answer in [0, K-1]
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (u >= lv.edges[mid] - e0) lo = mid;
            else hi = mid - 1;
        }
        return segment_value(lv, lo, inverse);
    };
    double origin = fmod(e0, period);
    if (origin < 0) origin += period;
    const double offset = origin / period + 0.5 / nfft;
    const double scale = 1.0 / nfft;
    if ((nfft & (nfft - 1)) == 0) {
        cd* a = buf;
        cd* b = buf + nfft;
        for (int j = threadIdx.x; j < nfft; j += NT) a[j] = sample(j);
        __syncthreads();
        // radix-2 Stockham: stage Ns combines pairs of length-Ns transforms with twiddles exp(-i pi k / Ns)
        const int half = nfft >> 1;
        for (int Ns = 1; Ns < nfft; Ns <<= 1) {
            for (int j = threadIdx.x; j < half; j += NT) {
                const int k = j & (Ns - 1);
                double sw, cw;
                sincospi(-static_cast<double>(k) / Ns, &sw, &cw);
                const cd v0 = a[j], v1 = a[j + half] * cd{cw, sw};
                const int d = (j / Ns) * 2 * Ns + k;
                b[d] = v0 + v1;
                b[d + Ns] = v0 - v1;
            }
            __syncthreads();
            cd* t = a;
            a = b;
            b = t;
        }
        for (int hh = threadIdx.x; hh <= 2 * hmax; hh += NT) {
            const int h = hh - hmax;
            const int bin = h < 0 ? nfft + h : h;
            const double arg = -2.0 * kPi * h * offset;
            double sa, ca;
            sincos(arg, &sa, &ca);
            out[hh] = (scale * a[bin]) * cd{ca, sa};
        }
    } else {
        // X_h = sum_j x_j exp(-i 2 pi j h / nfft), the exponent reduced exactly (j h mod nfft) before sincospi
        for (int hh = threadIdx.x; hh <= 2 * hmax; hh += NT) {
            const int h = hh - hmax;
            const long long hm = ((long long)h % nfft + nfft) % nfft;
            cd sum{0.0, 0.0};
            for (int j = 0; j < nfft; ++j) {
                const long long r = (hm * j) % nfft;
                double sw, cw;
                sincospi(-2.0 * static_cast<double>(r) / nfft, &sw, &cw);
                sum = sum + sample(j) * cd{cw, sw};
            }
            const double arg = -2.0 * kPi * h * offset;
            double sa, ca;
            sincos(arg, &sa, &ca);
            out[hh] = (scale * sum) * cd{ca, sa};
        }
    }
    __syncthreads();
    return 0;
}

// The layer's Fourier coefficients by its method (convolution_matrix, rcwa1d.hpp:391-407): analytic, or
// FFT-sampled for a grating with FourierMethod::FFT (fftbuf: 2 nfft scratch).  Returns 1 on failure.
DEV int layer_coefficients(const LayerView& lv, double period, int M, bool inverse, cd* out, cd* fftbuf) {
    if (!lv.homogeneous && lv.method == 1)
        return fft_coefficients(lv, period, 2 * M, inverse, fft_samples_for(M, lv.samples), fftbuf, out);
    fourier_coefficients(lv, period, 2 * M, inverse, out);
    return 0;
}

// [[f]](m, p) = c_{m-p}; c holds 2N-1 coefficients with c_0 at index N-1.
DEV void toeplitz(const cd* c, int N, cd* S) {
    wg_map(N * N, [&](int e) { return ldc(c + (e % N) - (e / N) + N - 1); }, [&](int e, cd v) { stc(S + e, v); });
}

// Re q >= 0; propagating (Re q ~ 0) takes Im q <= 0 (rcwa1d.hpp:514-518).
DEV cd branch_q(cd q) {
    if (q.x < 0.0) q = -q;
    if (q.x < 1e-10 && q.y > 0.0) q = -q;
    return q;
}

// Homogeneous medium: kz = sqrt(n^2 - kx^2) (Im >= 0; Re >= 0 when Im ~ 0), q = -i kz,
// p = q (TE) or q / n^2 (TM); finite layers use q_basis = 1 for an exact zero root (rcwa1d.hpp:522-544).
DEV void analytic_modes_1d(cd n, const cd* kx, int N, int pol, bool fzb, cd* q, cd* p) {
    const cd n2 = n * n;
    for (int j = threadIdx.x; j < N; j += NT) {
        cd kz = csqrt(n2 - cd{kx[j].x * kx[j].x, 0.0});
        if (kz.y < 0.0) kz = -kz;
        if (fabs(kz.y) < 1e-12 && kz.x < 0.0) kz = -kz;
        const cd qq{kz.y, -kz.x};
        q[j] = qq;
        const cd basis = (fzb && is_zero(qq)) ? cd{1.0, 0.0} : qq;
        p[j] = pol == 0 ? basis : cdiv(basis, n2);
    }
}

// ------------------------------------------------------------------------------------------------
// Bump allocator over one instance's workspace: temporaries from the low end (released per
// S-matrix block), kept results from the high end.  Requests that do not fit set the workspace
// failure and return the start of the region, so every access stays inside the instance's slab.
struct Bump {
    cd* ws;
    int64_t lo, hi, lo0;
    int overflow;
    DEV Bump() : ws(nullptr), lo(0), hi(0), lo0(0), overflow(0) {}
    DEV Bump(cd* w, int64_t l, int64_t h) : ws(w), lo(l), hi(h), lo0(l), overflow(0) {}
    DEV cd* tmp(int64_t n) {
        n = (n + 15) & ~int64_t(15);
        if (lo + n > hi) {
            overflow = 1;
            return ws + lo0;
        }
        cd* r = ws + lo;
        lo += n;
        return r;
    }
    DEV cd* keep(int64_t n) {
        n = (n + 15) & ~int64_t(15);
        if (hi - n < lo) {
            overflow = 1;
            return ws + lo0;
        }
        hi -= n;
        return ws + hi;
    }
};

struct SmatCtx {
    int N;
    Bump bump;
    int fail;       // 1 numerical, 2 workspace
    int n_solve = 0;        // dense linear solves (Result::n_dense_solve)
    bool diag = false;      // compute backward errors (Options::diagnostics)
    double max_berr = 0.0;  // largest backward error so far (Result::max_solve_backward_error)
    bool mixed = false;     // dense solves by FP32 LU + FP64 refinement (PLAN_MIXED)
    int n_refine = 0;       // FP32 correction steps over all refined solves
    int n_fallback = 0;     // refined solves that fell back to the FP64 LU
    double max_refine_res = 0.0;   // largest final ||b - A x|| / (||A|| ||x||) of a refined solve
    DEV void check() {
        if (bump.overflow) fail = 2;
    }
    // writes the refinement record (DIAG_WORDS slots 5..7) of one instance
    DEV void refine_record(double* dg) const {
        dg[5] = n_fallback;
        dg[6] = n_refine;
        dg[7] = max_refine_res;
    }
};

// max that keeps a NaN (a NaN residual or error marks a failed quantity; dt::max_keep_nan)
DEV double max_keep_nan(double a, double b) { return (isnan(a) || isnan(b)) ? __builtin_nan("") : fmax(a, b); }

// Backward error ||A X - B||_F / (||A||_F ||X||_F + ||B||_F) of a solve (solve_backward_error,
// lu.hpp:452-459).  A is n x n, X and B n x nrhs; R (n x nrhs) is scratch.
DEV double wg_backward_error(int n, const cd* A, const cd* X, const cd* B, int nrhs, cd* R) {
    const int64_t nn = (int64_t)n * n, nr = (int64_t)n * nrhs;
    wg_copy(n, nrhs, B, n, R, n);
    wg_gemm<GEMM_SUB, false>(n, nrhs, n, A, n, X, n, R, n);
    double sa = 0.0, sx = 0.0, sb = 0.0, sr = 0.0;
    for (int64_t e = threadIdx.x; e < nn; e += NT) sa += norm2(A[e]);
    for (int64_t e = threadIdx.x; e < nr; e += NT) {
        sx += norm2(X[e]);
        sb += norm2(B[e]);
        sr += norm2(R[e]);
    }
    sa = block_sum(sa);
    sx = block_sum(sx);
    sb = block_sum(sb);
    sr = block_sum(sr);
    const double denom = sqrt(sa) * sqrt(sx) + sqrt(sb), resid = sqrt(sr);
    if (denom == 0.0) return resid == 0.0 ? 0.0 : __builtin_inf();
    return resid / denom;
}

// Eigen-pair residual max_j ||A v_j - l_j v_j||_2 / (||A||_F ||v_j||_2) (eig_residual, eig.hpp:38-69):
// A and V n x n, w the n eigenvalues, AV (n x n) scratch.  A zero vector counts as +inf, a NaN as NaN.
DEV double wg_eig_residual(int n, const cd* A, const cd* V, const cd* w, cd* AV) {
    const int64_t nn = (int64_t)n * n;
    wg_gemm<GEMM_SET, false>(n, n, n, A, n, V, n, AV, n);
    double sa = 0.0;
    for (int64_t e = threadIdx.x; e < nn; e += NT) sa += norm2(A[e]);
    const double anorm = sqrt(block_sum(sa));
    double worst = 0.0;
    for (int j = wave_id(); j < n; j += NWAVE) {
        const cd* v = V + (int64_t)j * n;
        const cd* av = AV + (int64_t)j * n;
        const cd lam = w[j];
        double vs = 0.0, ds = 0.0;
        for (int i = lane_id(); i < n; i += 64) {
            vs += norm2(v[i]);
            ds += norm2(av[i] - lam * v[i]);
        }
        vs = sqrt(wave_sum(vs));
        ds = sqrt(wave_sum(ds));
        double value;
        if (vs == 0.0) value = __builtin_inf();
        else if (anorm == 0.0) value = ds == 0.0 ? 0.0 : __builtin_inf();
        else value = (ds / anorm) / vs;
        worst = max_keep_nan(worst, value);
    }
    // block max that keeps a NaN: NaN -> +inf marker through block_max, then back
    __shared__ int s_nan;
    if (threadIdx.x == 0) s_nan = 0;
    __syncthreads();
    if (isnan(worst)) s_nan = 1;
    const double m = block_max(isnan(worst) ? 0.0 : worst);
    const int any_nan = s_nan;
    __syncthreads();
    return any_nan ? __builtin_nan("") : m;
}

// zero / diagonal / dense operator of order N (rcwa1d.hpp:549-629)
enum { OP_ZERO = 0, OP_DIAG = 1, OP_DENSE = 2 };
struct DOp {
    int kind;
    cd* p;
};

// Modes of one region.  Homogeneous regions are diagonal: planar ones carry p (V = diag(p), W = I,
// w == nullptr); conical ones carry w and v in the per-order (s, t) basis (p holds v).  General
// regions carry dense W, V.  A layer propagated in its Schur basis carries the dense
// exp(-k0 d G) in X (conical degenerate layers, rcwa1d_conical.hpp:413-502).
struct Region {
    bool analytic;
    bool fzb;           // finite-layer zero-root basis (planar only)
    cd* q;
    cd* p;
    cd* w;
    cd* W;
    cd* V;
    cd* X;
};

struct SBlock {
    DOp Tdd, Rdu, Rud, Tuu;
};

DEV DOp op_zero() { return DOp{OP_ZERO, nullptr}; }

DEV DOp op_mul(SmatCtx& c, DOp a, DOp b) {
    const int N = c.N;
    if (a.kind == OP_ZERO || b.kind == OP_ZERO) return op_zero();
    if (a.kind == OP_DIAG && b.kind == OP_DIAG) {
        cd* v = c.bump.tmp(N);
        for (int i = threadIdx.x; i < N; i += NT) v[i] = a.p[i] * b.p[i];
        __syncthreads();
        return DOp{OP_DIAG, v};
    }
    cd* m = c.bump.tmp((int64_t)N * N);
    if (a.kind == OP_DIAG) wg_scale_rows(N, N, a.p, b.p, N, m, N);
    else if (b.kind == OP_DIAG) wg_scale_cols(N, N, a.p, N, b.p, m, N);
    else wg_gemm<GEMM_SET, false>(N, N, N, a.p, N, b.p, N, m, N);
    return DOp{OP_DENSE, m};
}

DEV DOp op_add(SmatCtx& c, DOp a, DOp b) {
    const int N = c.N;
    if (b.kind == OP_ZERO) return a;
    if (a.kind == OP_ZERO) return b;
    if (a.kind == OP_DIAG && b.kind == OP_DIAG) {
        cd* v = c.bump.tmp(N);
        for (int i = threadIdx.x; i < N; i += NT) v[i] = a.p[i] + b.p[i];
        __syncthreads();
        return DOp{OP_DIAG, v};
    }
    const int64_t NN = (int64_t)N * N;
    cd* m = c.bump.tmp(NN);
    if (a.kind == OP_DENSE && b.kind == OP_DENSE)
        wg_map((int)NN, [&](int e) { return ldc(a.p + e) + ldc(b.p + e); }, [&](int e, cd v) { stc(m + e, v); });
    else if (a.kind == OP_DENSE)
        wg_map((int)NN, [&](int e) { const int i = e % N; return i == e / N ? ldc(a.p + e) + ldc(b.p + i) : ldc(a.p + e); },
               [&](int e, cd v) { stc(m + e, v); });
    else
        wg_map((int)NN, [&](int e) { const int i = e % N; return i == e / N ? ldc(b.p + e) + ldc(a.p + i) : ldc(b.p + e); },
               [&](int e, cd v) { stc(m + e, v); });
    return DOp{OP_DENSE, m};
}

// y = a x (y must not alias x)
DEV void op_apply(SmatCtx& c, DOp a, const cd* x, cd* y) {
    const int N = c.N;
    for (int i = threadIdx.x; i < N; i += NT) {
        if (a.kind == OP_DENSE) {
            cd s{0.0, 0.0};
            for (int j = 0; j < N; ++j) s = s + a.p[i + (int64_t)j * N] * x[j];
            y[i] = s;
        } else if (a.kind == OP_DIAG) {
            y[i] = a.p[i] * x[i];
        } else {
            y[i] = cd{0.0, 0.0};
        }
    }
    __syncthreads();
}

// Y = a X for nrhs columns (Y must not alias X)
DEV void op_apply_cols(SmatCtx& c, DOp a, const cd* x, cd* y, int nrhs) {
    const int N = c.N;
    if (a.kind == OP_DENSE && nrhs > 1) {
        wg_gemm<GEMM_SET, false>(N, nrhs, N, a.p, N, x, N, y, N);
        return;
    }
    for (int k = 0; k < nrhs; ++k) {
        const cd* xk = x + (int64_t)k * N;
        cd* yk = y + (int64_t)k * N;
        for (int i = threadIdx.x; i < N; i += NT) {
            if (a.kind == OP_DENSE) {
                cd s{0.0, 0.0};
                for (int j = 0; j < N; ++j) s = s + a.p[i + (int64_t)j * N] * xk[j];
                yk[i] = s;
            } else if (a.kind == OP_DIAG) {
                yk[i] = a.p[i] * xk[i];
            } else {
                yk[i] = cd{0.0, 0.0};
            }
        }
    }
    __syncthreads();
}

DEV cd* op_to_dense(SmatCtx& c, DOp a) {
    const int N = c.N;
    cd* m = c.bump.tmp((int64_t)N * N);
    if (a.kind == OP_DENSE) wg_copy(N, N, a.p, N, m, N);
    else if (a.kind == OP_DIAG) wg_diag(N, a.p, m, N);
    else wg_fill((int64_t)N * N, m, cd{0.0, 0.0});
    return m;
}

#if SRCWA_MIXED_PRECISION
// The mixed-precision form of one dense solve: B <- A^{-1} B by FP32 LU and FP64 iterative refinement
// (wg_refine_solve), A preserved; all of its scratch is released on return.
DEV void mixed_dense_solve(SmatCtx& c, const cd* A, cd* B, int nrhs) {
    const int N = c.N;
    const int64_t mark = c.bump.lo;
    cd* lu64 = c.bump.tmp((int64_t)N * N);
    int* piv = reinterpret_cast<int*>(c.bump.tmp(N));
    cd* scr = c.bump.tmp(refine_words(N, nrhs));
    cd* b0 = c.diag ? c.bump.tmp((int64_t)N * nrhs) : nullptr;
    cd* r = c.diag ? c.bump.tmp((int64_t)N * nrhs) : nullptr;
    c.check();
    if (c.fail) return;
    if (c.diag) wg_copy(N, nrhs, B, N, b0, N);
    const RefineStats st = wg_refine_solve(N, A, B, nrhs, scr, lu64, piv);
    if (st.singular) {
        c.fail = 1;
        return;
    }
    c.n_refine += st.iters;
    c.n_fallback += st.fallback;
    c.max_refine_res = max_keep_nan(c.max_refine_res, st.resid);
    if (c.diag) c.max_berr = max_keep_nan(c.max_berr, wg_backward_error(N, A, B, b0, nrhs, r));
    c.bump.lo = mark;
}
#endif

// B <- A^{-1} B (A is preserved).  Singular -> numerical failure.  Every call is one dense solve of
// the reference (linear_solve, rcwa1d.hpp:248-281); with diagnostics its backward error is taken from
// copies of A and B released right after.
DEV void dense_solve(SmatCtx& c, const cd* A, cd* B, int nrhs) {
    const int N = c.N;
    ++c.n_solve;
#if SRCWA_MIXED_PRECISION
    if (c.mixed) {
        mixed_dense_solve(c, A, B, nrhs);
        return;
    }
#endif
    cd* lu = c.bump.tmp((int64_t)N * N);
    int* piv = reinterpret_cast<int*>(c.bump.tmp(N));
    c.check();
    if (c.fail) return;
    const int64_t mark = c.bump.lo;
    cd* b0 = nullptr;
    if (c.diag) {
        b0 = c.bump.tmp((int64_t)N * nrhs);
        c.check();
        if (c.fail) return;
        wg_copy(N, nrhs, B, N, b0, N);
    }
    wg_copy(N, N, A, N, lu, N);
    if (wg_lu(N, lu, N, piv)) {
        c.fail = 1;
        return;
    }
    wg_lu_solve(N, lu, N, piv, B, N, nrhs);
    if (c.diag) {
        cd* r = c.bump.tmp((int64_t)N * nrhs);
        c.check();
        if (c.fail) return;
        c.max_berr = max_keep_nan(c.max_berr, wg_backward_error(N, A, B, b0, nrhs, r));
        c.bump.lo = mark;
    }
}
// B <- A^{-1} B, destroying A (the mixed-precision form leaves A as it was).
DEV void dense_solve_inplace(SmatCtx& c, cd* A, cd* B, int nrhs) {
    const int N = c.N;
    ++c.n_solve;
#if SRCWA_MIXED_PRECISION
    if (c.mixed) {
        mixed_dense_solve(c, A, B, nrhs);
        return;
    }
#endif
    int* piv = reinterpret_cast<int*>(c.bump.tmp(N));
    c.check();
    if (c.fail) return;
    const int64_t mark = c.bump.lo;
    cd *a0 = nullptr, *b0 = nullptr;
    if (c.diag) {
        a0 = c.bump.tmp((int64_t)N * N);
        b0 = c.bump.tmp((int64_t)N * nrhs);
        c.check();
        if (c.fail) return;
        wg_copy(N, N, A, N, a0, N);
        wg_copy(N, nrhs, B, N, b0, N);
    }
    if (wg_lu(N, A, N, piv)) {
        c.fail = 1;
        return;
    }
    wg_lu_solve(N, A, N, piv, B, N, nrhs);
    if (c.diag) {
        cd* r = c.bump.tmp((int64_t)N * nrhs);
        c.check();
        if (c.fail) return;
        c.max_berr = max_keep_nan(c.max_berr, wg_backward_error(N, a0, B, b0, nrhs, r));
        c.bump.lo = mark;
    }
}

// Both sides homogeneous: per-order Fresnel blocks (rcwa1d.hpp:649-669).
DEV SBlock diagonal_interface(SmatCtx& c, const Region& up, const Region& low) {
    const int N = c.N;
    cd* tdd = c.bump.tmp(N);
    cd* rdu = c.bump.tmp(N);
    cd* rud = c.bump.tmp(N);
    cd* tuu = c.bump.tmp(N);
    if (up.w == nullptr) {
        for (int j = threadIdx.x; j < N; j += NT) {
            const cd p1 = up.p[j], p2 = low.p[j];
            const cd inv = cdiv(cd{1.0, 0.0}, p1 + p2);
            tdd[j] = (2.0 * p1) * inv;
            rdu[j] = (p2 - p1) * inv;
            rud[j] = (p1 - p2) * inv;
            tuu[j] = (2.0 * p2) * inv;
        }
    } else {
        // (w, v) form, finite when one side grazes (rcwa1d_conical.hpp:507-527)
        for (int j = threadIdx.x; j < N; j += NT) {
            const cd w1 = up.w[j], v1 = up.p[j], w2 = low.w[j], v2 = low.p[j];
            const cd a = w1 * v2, b = v1 * w2;
            const cd inv = cdiv(cd{1.0, 0.0}, a + b);
            tdd[j] = (2.0 * (w1 * v1)) * inv;
            rdu[j] = (a - b) * inv;
            rud[j] = (b - a) * inv;
            tuu[j] = (2.0 * (w2 * v2)) * inv;
        }
    }
    __syncthreads();
    return SBlock{DOp{OP_DIAG, tdd}, DOp{OP_DIAG, rdu}, DOp{OP_DIAG, rud}, DOp{OP_DIAG, tuu}};
}

// T+ and T- (rcwa1d.hpp:672-692)
DEV void interface_matching(SmatCtx& c, const Region& up, const Region& low, cd*& tp, cd*& tm) {
    const int N = c.N;
    const int64_t NN = (int64_t)N * N;
    cd* mm = c.bump.tmp(NN);
    cd* nn = c.bump.tmp(NN);
    c.check();
    if (c.fail) return;
    if (low.analytic && low.w != nullptr) {
        // W2 = diag(w2), V2 = diag(v2): row scaling; a zero (grazing order) makes the interface singular
        int zero = 0;
        for (int j = threadIdx.x; j < N; j += NT) zero |= is_zero(low.w[j]) || is_zero(low.p[j]);
        if (block_max_int(zero)) {
            c.fail = 1;
            return;
        }
        wg_map((int)NN, [&](int e) { return cdiv(cd{1.0, 0.0}, ldc(low.w + e % N)) * ldc(up.W + e); },
               [&](int e, cd v) { stc(mm + e, v); });
        wg_map((int)NN, [&](int e) { return cdiv(cd{1.0, 0.0}, ldc(low.p + e % N)) * ldc(up.V + e); },
               [&](int e, cd v) { stc(nn + e, v); });
    } else if (low.analytic) {
        wg_copy(N, N, up.W, N, mm, N);
        wg_map((int)NN, [&](int e) { return cdiv(cd{1.0, 0.0}, ldc(low.p + e % N)) * ldc(up.V + e); },
               [&](int e, cd v) { stc(nn + e, v); });
    } else if (up.analytic) {
        if (up.w != nullptr) wg_diag(N, up.w, mm, N);
        else wg_identity(N, mm, N);
        dense_solve(c, low.W, mm, N);
        wg_diag(N, up.p, nn, N);
        dense_solve(c, low.V, nn, N);
    } else {
        wg_copy(N, N, up.W, N, mm, N);
        dense_solve(c, low.W, mm, N);
        wg_copy(N, N, up.V, N, nn, N);
        dense_solve(c, low.V, nn, N);
    }
    if (c.fail) return;
    struct Pair {
        cd p, m;
    };
    wg_map((int)NN,
           [&](int e) {
               const cd a = ldc(mm + e), b = ldc(nn + e);
               return Pair{0.5 * (a + b), 0.5 * (a - b)};
           },
           [&](int e, Pair v) {
               stc(mm + e, v.p);
               stc(nn + e, v.m);
           });
    tp = mm;
    tm = nn;
}

DEV SBlock interface_block(SmatCtx& c, const Region& up, const Region& low) {
    if (up.analytic && low.analytic) return diagonal_interface(c, up, low);
    const int N = c.N;
    const int64_t NN = (int64_t)N * N;
    cd *tp = nullptr, *tm = nullptr;
    interface_matching(c, up, low, tp, tm);
    SBlock s{op_zero(), op_zero(), op_zero(), op_zero()};
    if (c.fail) return s;
    cd* tpi = c.bump.tmp(NN);
    cd* rdu = c.bump.tmp(NN);
    cd* rud = c.bump.tmp(NN);
    c.check();
    if (c.fail) return s;
    wg_identity(N, tpi, N);
    dense_solve(c, tp, tpi, N);
    if (c.fail) return s;
    wg_gemm<GEMM_SET, false>(N, N, N, tm, N, tpi, N, rdu, N);       // Rdu = T- T+^-1
    wg_gemm<GEMM_SUB, false>(N, N, N, rdu, N, tm, N, tp, N);        // Tdd = T+ - Rdu T-
    wg_fill(NN, rud, cd{0.0, 0.0});
    wg_gemm<GEMM_SUB, false>(N, N, N, tpi, N, tm, N, rud, N);       // Rud = -T+^-1 T-
    return SBlock{DOp{OP_DENSE, tp}, DOp{OP_DENSE, rdu}, DOp{OP_DENSE, rud}, DOp{OP_DENSE, tpi}};
}

// exp(-k0 q d); finite-layer exact zero roots use 1/(1+h/2) and (h/2)/(1+h/2) (rcwa1d.hpp:751-784).
DEV SBlock propagation_block(SmatCtx& c, const cd* q, double d, double k0, bool fzb) {
    const int N = c.N;
    cd* x = c.bump.tmp(N);
    cd* refl = c.bump.tmp(N);
    int zero = 0, bad = 0;
    const double h = k0 * d;
    for (int j = threadIdx.x; j < N; j += NT) {
        if (fzb && is_zero(q[j])) {
            if (!isfinite(h)) bad = 1;
            const double den = 1.0 + 0.5 * h;
            x[j] = cd{1.0 / den, 0.0};
            refl[j] = cd{(0.5 * h) / den, 0.0};
            zero = 1;
        } else {
            x[j] = cexp((-k0 * d) * q[j]);
            refl[j] = cd{0.0, 0.0};
        }
    }
    zero = block_max_int(zero);
    bad = block_max_int(bad);
    if (bad) c.fail = 1;
    if (zero) return SBlock{DOp{OP_DIAG, x}, DOp{OP_DIAG, refl}, DOp{OP_DIAG, refl}, DOp{OP_DIAG, x}};
    return SBlock{DOp{OP_DIAG, x}, op_zero(), op_zero(), DOp{OP_DIAG, x}};
}

// g = (I - Rdu r)^-1 Tdd, r' = Rud + Tuu r g (rcwa1d.hpp:794-808)
DEV void stack_on(SmatCtx& c, const SBlock& s, DOp r, DOp& g, DOp& rn) {
    const int N = c.N;
    DOp k = op_mul(c, s.Rdu, r);
    if (k.kind == OP_ZERO) {
        g = s.Tdd;
    } else if (k.kind == OP_DIAG) {
        for (int i = threadIdx.x; i < N; i += NT) k.p[i] = cdiv(cd{1.0, 0.0}, cd{1.0, 0.0} - k.p[i]);
        __syncthreads();
        g = op_mul(c, k, s.Tdd);
    } else {
        const int64_t NN = (int64_t)N * N;
        wg_map((int)NN,
               [&](int e) {
                   cd v = -ldc(k.p + e);
                   if (e % N == e / N) v = v + cd{1.0, 0.0};
                   return v;
               },
               [&](int e, cd v) { stc(k.p + e, v); });
        cd* rhs = op_to_dense(c, s.Tdd);
        c.check();
        if (c.fail) return;
        dense_solve_inplace(c, k.p, rhs, N);
        g = DOp{OP_DENSE, rhs};
    }
    c.check();
    if (c.fail) return;
    const DOp rg = op_mul(c, r, g);
    const DOp t = op_mul(c, s.Tuu, rg);
    rn = op_add(c, s.Rud, t);
    c.check();
}

// Persist an operator into kept storage.
DEV DOp op_keep(SmatCtx& c, DOp a, cd* slot) {
    const int N = c.N;
    if (a.kind == OP_ZERO) return a;
    const int64_t cnt = a.kind == OP_DIAG ? N : (int64_t)N * N;
    if (a.p != slot) wg_map((int)cnt, [&](int e) { return ldc(a.p + e); }, [&](int e, cd v) { stc(slot + e, v); });
    return DOp{a.kind, slot};
}

// The whole recursion (rcwa1d.hpp:963-1034; conical rcwa1d_conical.hpp:859-919): bottom-up seams,
// the top block restricted to the incident columns cols[0..nrhs), the downward sweep.
// region(r) returns the modes of region r (0 incidence, j+1 layer j, L+1 exit).  Writes the
// reflected (up) and transmitted (down) amplitudes, N x nrhs each.  On return c.fail says whether
// the instance failed.
//
// Near-field capture (rcwa1d.hpp:968-1023, rcwa1d_conical.hpp:859-919): with fields != null the reflection
// below each layer is kept as the sweep passes it, and the downward sweep writes each layer's amplitudes
// (N x nrhs each, column-major), a+ at the layer top (fields + j N nrhs) and a- = r_below (the amplitudes
// at the layer bottom) (fields + (L + j) N nrhs).  pure: every layer is a uniform film; then only the rows
// of diffraction order keep_order (row i is order i mod orders) are kept, as the reference clears the
// others (they are not excited).
template <class RegionFn>
DEV void run_recursion(SmatCtx& c, int L, double k0, const double* thick, RegionFn region, int nrhs, const int* cols,
                       cd* up_out, cd* down_out, cd* fields = nullptr, bool pure = false, int keep_order = 0,
                       int orders = 1) {
    __shared__ int g_kind[SMAT_MAX_BLOCKS];
    __shared__ cd* g_ptr[SMAT_MAX_BLOCKS];
    __shared__ int rb_kind[SMAT_MAX_LAYERS];
    __shared__ cd* rb_ptr[SMAT_MAX_LAYERS];
    const int N = c.N;
    const int64_t NN = (int64_t)N * N;
    const int nb = 2 * L + 1;
    cd* rslot[2] = {c.bump.keep(NN), c.bump.keep(NN)};
    int rs = 0;
    DOp r = op_zero();
    c.check();
    for (int b = nb - 1; b >= 1 && !c.fail; --b) {
        const int64_t mark = c.bump.lo;
        const int j = b / 2;
        SBlock blk;
        if (b % 2 == 0) {
            blk = interface_block(c, region(j), region(j + 1));
        } else {
            const Region g = region(j + 1);
            if (g.X != nullptr) blk = SBlock{DOp{OP_DENSE, g.X}, op_zero(), op_zero(), DOp{OP_DENSE, g.X}};
            else blk = propagation_block(c, g.q, thick[j], k0, g.fzb);
        }
        c.check();
        if (c.fail) break;
        if (fields && (b & 1)) {
            // the reflection below layer j (rcwa1d.hpp:978), kept for its a- (slots from the kept end)
            DOp kept = r;
            if (r.kind != OP_ZERO) kept = op_keep(c, r, c.bump.keep(r.kind == OP_DIAG ? N : NN));
            c.check();
            if (c.fail) break;
            if (threadIdx.x == 0) {
                rb_kind[j] = kept.kind;
                rb_ptr[j] = kept.p;
            }
        }
        DOp g, rn;
        stack_on(c, blk, r, g, rn);
        if (c.fail) break;
        cd* gslot = c.bump.keep(g.kind == OP_DIAG ? N : NN);
        c.check();
        if (c.fail) break;
        g = op_keep(c, g, gslot);
        if (threadIdx.x == 0) {
            g_kind[b] = g.kind;
            g_ptr[b] = g.p;
        }
        r = op_keep(c, rn, rslot[rs ^ 1]);
        rs ^= 1;
        c.bump.lo = mark;
        __syncthreads();
    }
    if (c.fail) return;

    // top block: only the incident columns of Tdd and Rud (rcwa1d.hpp:715-746, rcwa1d_conical.hpp:765-795)
    const Region up = region(0), low = region(1);
    const int64_t NR = (int64_t)N * nrhs;
    cd* tdd = c.bump.tmp(NR);
    cd* rud = c.bump.tmp(NR);
    DOp Rdu, Tuu;
    if (up.analytic && low.analytic) {
        const SBlock s = diagonal_interface(c, up, low);
        for (int64_t e = threadIdx.x; e < NR; e += NT) {
            const int j = (int)(e % N), k = (int)(e / N);
            tdd[e] = j == cols[k] ? s.Tdd.p[j] : cd{0.0, 0.0};
            rud[e] = j == cols[k] ? s.Rud.p[j] : cd{0.0, 0.0};
        }
        __syncthreads();
        Rdu = s.Rdu;
        Tuu = s.Tuu;
    } else {
        cd *tp = nullptr, *tm = nullptr;
        interface_matching(c, up, low, tp, tm);
        if (c.fail) return;
        cd* tpi = c.bump.tmp(NN);
        cd* rdu = c.bump.tmp(NN);
        cd* tme = c.bump.tmp(NR);
        c.check();
        if (c.fail) return;
        wg_identity(N, tpi, N);
        dense_solve(c, tp, tpi, N);
        if (c.fail) return;
        wg_gemm<GEMM_SET, false>(N, N, N, tm, N, tpi, N, rdu, N);
        for (int64_t e = threadIdx.x; e < NR; e += NT) {
            const int j = (int)(e % N), k = (int)(e / N);
            tme[e] = tm[j + (int64_t)cols[k] * N];
            tdd[e] = tp[j + (int64_t)cols[k] * N];
        }
        __syncthreads();
        wg_gemm<GEMM_SUB, false>(N, nrhs, N, rdu, N, tme, N, tdd, N);    // Tdd e = T+ e - Rdu (T- e)
        wg_fill(NR, rud, cd{0.0, 0.0});
        wg_gemm<GEMM_SUB, false>(N, nrhs, N, tpi, N, tme, N, rud, N);    // Rud e = -T+^-1 (T- e)
        Rdu = DOp{OP_DENSE, rdu};
        Tuu = DOp{OP_DENSE, tpi};
    }
    DOp k = op_mul(c, Rdu, r);
    cd* down = c.bump.tmp(NR);
    cd* tmpv = c.bump.tmp(NR);
    c.check();
    if (c.fail) return;
    for (int64_t e = threadIdx.x; e < NR; e += NT) down[e] = tdd[e];
    __syncthreads();
    if (k.kind == OP_DIAG) {
        for (int64_t e = threadIdx.x; e < NR; e += NT) {
            const int j = (int)(e % N);
            down[e] = cdiv(down[e], cd{1.0, 0.0} - k.p[j]);
        }
        __syncthreads();
    } else if (k.kind == OP_DENSE) {
        wg_map((int)NN,
               [&](int e) {
                   cd v = -ldc(k.p + e);
                   if (e % N == e / N) v = v + cd{1.0, 0.0};
                   return v;
               },
               [&](int e, cd v) { stc(k.p + e, v); });
        dense_solve_inplace(c, k.p, down, nrhs);
        if (c.fail) return;
    }
    op_apply_cols(c, r, down, tmpv, nrhs);
    op_apply_cols(c, Tuu, tmpv, up_out, nrhs);
    for (int64_t e = threadIdx.x; e < NR; e += NT) up_out[e] = rud[e] + up_out[e];
    __syncthreads();
    cd* bottom = fields ? c.bump.tmp(NR) : nullptr;
    c.check();
    if (c.fail) return;
    // pure films: only the incident order is excited (rcwa1d.hpp:1011-1020)
    auto excited = [&](int64_t e) { return !pure || (int)(e % N) % orders == keep_order; };
    for (int b = 1; b < nb; ++b) {
        const int j = b / 2;
        if (fields && (b & 1))
            for (int64_t e = threadIdx.x; e < NR; e += NT) fields[j * NR + e] = excited(e) ? down[e] : cd{0.0, 0.0};
        op_apply_cols(c, DOp{g_kind[b], g_ptr[b]}, down, tmpv, nrhs);
        cd* sw = down;
        down = tmpv;
        tmpv = sw;
        if (fields && (b & 1)) {
            for (int64_t e = threadIdx.x; e < NR; e += NT) bottom[e] = excited(e) ? down[e] : cd{0.0, 0.0};
            __syncthreads();
            cd* am = fields + (L + j) * NR;
            op_apply_cols(c, DOp{rb_kind[j], rb_ptr[j]}, bottom, am, nrhs);
            if (pure) {
                for (int64_t e = threadIdx.x; e < NR; e += NT)
                    if (!excited(e)) am[e] = cd{0.0, 0.0};
                __syncthreads();
            }
        }
    }
    for (int64_t e = threadIdx.x; e < NR; e += NT) down_out[e] = down[e];
    __syncthreads();
}
