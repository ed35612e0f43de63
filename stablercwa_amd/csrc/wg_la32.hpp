// Complex-FP32 LU and substitution for the mixed-precision star products, and the FP64 iterative
// refinement that wins the FP64 answer back (the north-star precision form; not in the reference, whose
// LU is FP64 only, lu.hpp:267-443).
//
// Same conventions as wg_la.hpp: one whole 256-thread workgroup per matrix, uniform arguments, every
// routine ends with a barrier.  Matrices are column-major complex float (`cf`, interleaved re, im).
//
//   wg_lu32          P A = L U in FP32, partial pivoting by largest modulus (first maximum wins),
//                    32-column panels, trailing update C -= A B on MFMA v_mfma_f32_16x16x4f32
//   wg_lu_solve32    Y <- U^{-1} L^{-1} Y for a right-hand side already in pivot order (blocked
//                    substitution: diagonal blocks in registers, off-diagonal updates on MFMA)
//   wg_refine_solve  X = A^{-1} B to FP64 accuracy: FP32 factors + FP64 residual r = B - A X (FP64 MFMA)
//                    + FP32 corrections, per-column stopping test of LAPACK's zcgesv; a stall or
//                    an FP32 breakdown falls back to the FP64 LU
#pragma once

#include "batch.hpp"
#include "wg_la.hpp"

struct __attribute__((aligned(8))) cf {
    float x, y;
};
typedef float f4 __attribute__((ext_vector_type(4)));

DEV cf ldf(const cf* p) { return cf{p->x, p->y}; }
DEV void stf(cf* p, cf a) {
    p->x = a.x;
    p->y = a.y;
}
DEV cf fmul(cf a, cf b) { return cf{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
DEV cf fsub(cf a, cf b) { return cf{a.x - b.x, a.y - b.y}; }
// a - b c
DEV cf fsubmul(cf a, cf b, cf c) { return cf{a.x - (b.x * c.x - b.y * c.y), a.y - (b.x * c.y + b.y * c.x)}; }
DEV cf to_cf(cd a) { return cf{static_cast<float>(a.x), static_cast<float>(a.y)}; }
DEV cd to_cd(cf a) { return cd{static_cast<double>(a.x), static_cast<double>(a.y)}; }
// 1 / a by Smith's method
DEV cf frecip(cf a) {
    if (fabsf(a.x) >= fabsf(a.y)) {
        const float r = a.y / a.x, den = a.x + a.y * r;
        return cf{1.0f / den, -r / den};
    }
    const float r = a.x / a.y, den = a.x * r + a.y;
    return cf{r / den, -1.0f / den};
}

constexpr int LU32_NB = 32;      // FP32 LU / substitution block width

// C (m x n) -= A (m x k) B (k x n), complex FP32.  Each wave owns 32x32 output tiles (2x2 blocks of
// 16x16), a complex product being four real MFMAs per block.  v_mfma_f32_16x16x4f32 operands (per lane
// l): A[l%16][l/16], B[l/16][l%16]; accumulator element r of lane l is D[4 (l/16) + r][l%16].
DEV void wg_gemm32_sub(int m, int n, int k, const cf* __restrict__ A, int lda, const cf* __restrict__ B, int ldb,
                       cf* C, int ldc) {
    const int lane = lane_id();
    const int li = lane & 15, lk = lane >> 4;
    const int tm = (m + 31) >> 5, tn = (n + 31) >> 5, tiles = tm * tn;
    for (int t = wave_id(); t < tiles; t += NWAVE) {
        const int r0 = (t % tm) << 5, c0 = (t / tm) << 5;
        f4 re[2][2], im[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                re[i][j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
                im[i][j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
            }
        const int ra[2] = {r0 + li, r0 + 16 + li};
        const int cb[2] = {c0 + li, c0 + 16 + li};
        auto fetch = [&](int k0, cf* a, cf* b) {
            const int kk = k0 + lk;
            const bool kin = kk < k;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a[i] = (kin && ra[i] < m) ? ldf(A + ra[i] + (long)kk * lda) : cf{0.0f, 0.0f};
                b[i] = (kin && cb[i] < n) ? ldf(B + kk + (long)cb[i] * ldb) : cf{0.0f, 0.0f};
            }
        };
        cf a0[2], b0[2], a1[2], b1[2];
        fetch(0, a0, b0);
        fetch(4, a1, b1);
        for (int k0 = 0; k0 < k; k0 += 4) {
            cf a[2] = {a0[0], a0[1]}, b[2] = {b0[0], b0[1]};
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a0[i] = a1[i];
                b0[i] = b1[i];
            }
            if (k0 + 8 < k) fetch(k0 + 8, a1, b1);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    re[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].x, b[j].x, re[i][j], 0, 0, 0);
                    re[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(-a[i].y, b[j].y, re[i][j], 0, 0, 0);
                    im[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].x, b[j].y, im[i][j], 0, 0, 0);
                    im[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].y, b[j].x, im[i][j], 0, 0, 0);
                }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int col = c0 + 16 * j + li;
                if (col >= n) continue;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = r0 + 16 * i + 4 * lk + q;
                    if (row >= m) continue;
                    cf* p = C + row + (long)col * ldc;
                    const cf c = ldf(p);
                    stf(p, cf{c.x - re[i][j][q], c.y - im[i][j][q]});
                }
            }
    }
    __syncthreads();
}

// C (m x nrhs) -= A (m x k) B (k x nrhs) for the short k of the substitutions: MFMA tiles, or one VALU
// dot product per entry for fewer than 8 right-hand sides.
DEV void wg_update32(int m, int nrhs, int k, const cf* A, int lda, const cf* B, int ldb, cf* C, int ldc) {
    if (m <= 0) return;
    if (nrhs >= 8) {
        wg_gemm32_sub(m, nrhs, k, A, lda, B, ldb, C, ldc);
        return;
    }
    for (int e = threadIdx.x; e < m * nrhs; e += NT) {
        const int i = e % m, c = e / m;
        const cf* b = B + (long)c * ldb;
        cf s = ldf(C + i + (long)c * ldc);
        for (int kk = 0; kk < k; ++kk) s = fsubmul(s, ldf(A + i + (long)kk * lda), ldf(b + kk));
        stf(C + i + (long)c * ldc, s);
    }
    __syncthreads();
}

// P A = L U in place (FP32), right-looking with LU32_NB-column panels.  Row exchanges are applied to the
// whole row as each pivot is chosen (LAPACK getrf order), so L and U come out in pivot order and
// piv[k] is the row exchanged with row k at step k.  Returns 1 when a pivot is exactly zero or not
// finite (the FP32 factors are then unusable and the caller falls back to FP64).
DEV int wg_lu32(int n, cf* A, int lda, int* piv) {
    const int tid = threadIdx.x;
    for (int k0 = 0; k0 < n; k0 += LU32_NB) {
        const int nb = min(LU32_NB, n - k0), kend = k0 + nb;
        for (int k = k0; k < kend; ++k) {
            double best = -1.0;
            int bi = n;
            for (int i = k + tid; i < n; i += NT) {
                const cf a = ldf(A + i + (long)k * lda);
                const double m2 = static_cast<double>(a.x) * a.x + static_cast<double>(a.y) * a.y;
                if (m2 > best) {
                    best = m2;
                    bi = i;
                }
            }
            int p = block_argmax(best, bi);
            if (p >= n) p = k;                 // no finite modulus in the column: keep the row; the pivot test fails
            if (tid == 0) piv[k] = p;
            if (p != k)
                for (int j = tid; j < n; j += NT) {
                    cf* rk = A + k + (long)j * lda;
                    cf* rp = A + p + (long)j * lda;
                    const cf a = ldf(rk), b = ldf(rp);
                    stf(rk, b);
                    stf(rp, a);
                }
            __syncthreads();
            const cf pv = ldf(A + k + (long)k * lda);
            if ((pv.x == 0.0f && pv.y == 0.0f) || !isfinite(pv.x) || !isfinite(pv.y)) return 1;
            const cf inv = frecip(pv);
            for (int i = k + 1 + tid; i < n; i += NT) {
                cf* ai = A + i;
                const cf l = fmul(ldf(ai + (long)k * lda), inv);
                stf(ai + (long)k * lda, l);
                for (int j = k + 1; j < kend; ++j)
                    stf(ai + (long)j * lda, fsubmul(ldf(ai + (long)j * lda), l, ldf(A + k + (long)j * lda)));
            }
            __syncthreads();
        }
        if (kend >= n) break;
        // U12 = L11^{-1} A12, one thread per column right of the panel
        for (int j = kend + tid; j < n; j += NT) {
            cf* col = A + k0 + (long)j * lda;
            cf x[LU32_NB];
#pragma unroll
            for (int r = 0; r < LU32_NB; ++r)
                if (r < nb) x[r] = ldf(col + r);
#pragma unroll
            for (int r = 1; r < LU32_NB; ++r)
                if (r < nb) {
                    cf s = x[r];
#pragma unroll
                    for (int q = 0; q < LU32_NB; ++q)
                        if (q < r) s = fsubmul(s, ldf(A + k0 + r + (long)(k0 + q) * lda), x[q]);
                    x[r] = s;
                }
#pragma unroll
            for (int r = 0; r < LU32_NB; ++r)
                if (r < nb) stf(col + r, x[r]);
        }
        __syncthreads();
        wg_gemm32_sub(n - kend, n - kend, nb, A + kend + (long)k0 * lda, lda, A + k0 + (long)kend * lda, lda,
                      A + kend + (long)kend * lda, lda);
    }
    __syncthreads();
    return 0;
}

// Y <- U^{-1} L^{-1} Y from wg_lu32's factors; Y (n x nrhs, leading dimension ldy) must already be in
// pivot order.  Blocks of LU32_NB rows: the diagonal block solved in registers, one right-hand side per
// thread; the rest of the right-hand side updated by wg_update32.
DEV void wg_lu_solve32(int n, const cf* LU, int lda, cf* Y, int ldy, int nrhs) {
    const int tid = threadIdx.x;
    // forward: unit lower triangle
    for (int k0 = 0; k0 < n; k0 += LU32_NB) {
        const int nb = min(LU32_NB, n - k0);
        for (int j = tid; j < nrhs; j += NT) {
            cf* col = Y + k0 + (long)j * ldy;
            cf x[LU32_NB];
#pragma unroll
            for (int r = 0; r < LU32_NB; ++r)
                if (r < nb) x[r] = ldf(col + r);
#pragma unroll
            for (int r = 1; r < LU32_NB; ++r)
                if (r < nb) {
                    cf s = x[r];
#pragma unroll
                    for (int q = 0; q < LU32_NB; ++q)
                        if (q < r) s = fsubmul(s, ldf(LU + k0 + r + (long)(k0 + q) * lda), x[q]);
                    x[r] = s;
                }
#pragma unroll
            for (int r = 0; r < LU32_NB; ++r)
                if (r < nb) stf(col + r, x[r]);
        }
        __syncthreads();
        wg_update32(n - k0 - nb, nrhs, nb, LU + k0 + nb + (long)k0 * lda, lda, Y + k0, ldy, Y + k0 + nb, ldy);
    }
    // backward: upper triangle, last block first
    const int nblk = (n + LU32_NB - 1) / LU32_NB;
    for (int b = nblk - 1; b >= 0; --b) {
        const int k0 = b * LU32_NB, nb = min(LU32_NB, n - k0);
        for (int j = tid; j < nrhs; j += NT) {
            cf* col = Y + k0 + (long)j * ldy;
            cf x[LU32_NB];
#pragma unroll
            for (int r = 0; r < LU32_NB; ++r)
                if (r < nb) x[r] = ldf(col + r);
#pragma unroll
            for (int r = LU32_NB - 1; r >= 0; --r)
                if (r < nb) {
                    cf s = x[r];
#pragma unroll
                    for (int q = 0; q < LU32_NB; ++q)
                        if (q > r && q < nb) s = fsubmul(s, ldf(LU + k0 + r + (long)(k0 + q) * lda), x[q]);
                    x[r] = fmul(s, frecip(ldf(LU + k0 + r + (long)(k0 + r) * lda)));
                }
#pragma unroll
            for (int r = 0; r < LU32_NB; ++r)
                if (r < nb) stf(col + r, x[r]);
        }
        __syncthreads();
        wg_update32(k0, nrhs, nb, LU + (long)k0 * lda, lda, Y + k0, ldy, Y, ldy);
    }
}

// Row permutation of wg_lu32's exchanges as a gather list: perm[i] = the original row that ends in row i
// (the exchanges composed by one thread in order).  perm has n entries.
DEV void lu32_permutation(int n, const int* piv, int* perm) {
    for (int i = threadIdx.x; i < n; i += NT) perm[i] = i;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < n; ++k) {
            const int p = piv[k];
            if (p != k) {
                const int t = perm[k];
                perm[k] = perm[p];
                perm[p] = t;
            }
        }
    __syncthreads();
}

// Outcome of one refined solve.
struct RefineStats {
    int singular;       // 1: the FP64 fallback found an exactly zero pivot (numerical failure)
    int iters;          // FP32 correction steps after the first solve
    int fallback;       // 1: the FP64 LU was used instead
    double resid;       // max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf) at the end
};


// Largest number of FP32 correction steps before the solve falls back to FP64 (zcgesv's ITERMAX is 30;
// the corrections here contract by about kappa(A) * 2^-24 per step, so a solve that has not converged
// after a few steps is ill-conditioned for FP32 and the FP64 LU is cheaper than more steps).
constexpr int REFINE_MAX_ITERS = 6;

// X = A^{-1} B to FP64 accuracy; A (n x n) is read only, B (n x nrhs) is overwritten by X.  scratch holds
// refine_words(n, nrhs) words; lu64 (n x n complex double) and piv64 (n ints) are the FP64 fallback's
// space.  Stopping test per column j (LAPACK zcgesv): ||r_j||_inf <= ||x_j||_inf ||A||_inf eps sqrt(n),
// eps = 2^-53.  A correction that does not shrink by half, an FP32 pivot that is zero or not finite, a
// non-finite residual, or REFINE_MAX_ITERS steps without convergence: the FP64 LU solve instead.  Not
// inlined: one copy per kernel, however many solves call it.
__device__ __noinline__ RefineStats wg_refine_solve(int n, const cd* A, cd* B, int nrhs, cd* scratch, cd* lu64,
                                                    int* piv64) {
    RefineStats st{0, 0, 0, 0.0};
    const int tid = threadIdx.x;
    const int64_t nn = int64_t(n) * n, nr = int64_t(n) * nrhs;
    cd* base = scratch;
    cf* lu = reinterpret_cast<cf*>(base);
    base += (nn + 1) / 2 + 16;
    cd* x = base;
    base += nr + 16;
    cd* r = base;
    base += nr + 16;
    cf* y = reinterpret_cast<cf*>(base);
    base += (nr + 1) / 2 + 16;
    int* piv = reinterpret_cast<int*>(base);
    base += n / 4 + 16;
    int* perm = reinterpret_cast<int*>(base);
    __shared__ int s_flag;
    // ||A||_inf and the FP32 copy of A
    double rowmax = 0.0;
    for (int i = tid; i < n; i += NT) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += cabs(A[i + int64_t(j) * n]);
        rowmax = fmax(rowmax, s);
    }
    const double anorm = block_max(rowmax);
    wg_map(static_cast<int>(nn), [&](int e) { return to_cf(ldc(A + e)); }, [&](int e, cf v) { stf(lu + e, v); });
    int bad = wg_lu32(n, lu, n, piv) || !isfinite(anorm);
    if (!bad) {
        lu32_permutation(n, piv, perm);
        const double cte = anorm * ldexp(1.0, -53) * sqrt(static_cast<double>(n));
        // x = 0, r = B: the first step solves for the whole right-hand side
        wg_map(static_cast<int>(nr), [&](int e) { return ldc(B + e); }, [&](int e, cd v) { stc(r + e, v); stc(x + e, cd{0.0, 0.0}); });
        double prev = __builtin_inf();
        for (int it = 0;; ++it) {
            // y = P r in FP32, corrections d = U^{-1} L^{-1} y, x += d
            wg_map(static_cast<int>(nr),
                   [&](int e) { const int i = e % n, j = e / n; return to_cf(ldc(r + perm[i] + int64_t(j) * n)); },
                   [&](int e, cf v) { stf(y + e, v); });
            wg_lu_solve32(n, lu, n, y, n, nrhs);
            double dmax = 0.0;
            for (int64_t e = tid; e < nr; e += NT) {
                const cd d = to_cd(ldf(y + e));
                dmax = fmax(dmax, abs1(d));
                stc(x + e, ldc(x + e) + d);
            }
            dmax = block_max(dmax);      // also the barrier before the residual reads x
            // r = B - A x in FP64 (MFMA)
            wg_copy(n, nrhs, B, n, r, n);
            wg_gemm<GEMM_SUB, false>(n, nrhs, n, A, n, x, n, r, n);
            // per-column test; worst ratio ||r_j|| / (||A|| ||x_j||) for the record
            if (tid == 0) s_flag = 0;
            double worst = 0.0;
            __syncthreads();
            for (int j = wave_id(); j < nrhs; j += NWAVE) {
                double rm = 0.0, xm = 0.0;
                for (int i = lane_id(); i < n; i += 64) {
                    rm = fmax(rm, cabs(r[i + int64_t(j) * n]));
                    xm = fmax(xm, cabs(x[i + int64_t(j) * n]));
                }
                rm = wave_max(rm);
                xm = wave_max(xm);
                if (!(rm <= xm * cte) && lane_id() == 0) s_flag = 1;
                if (isnan(rm) || isnan(xm)) worst = __builtin_nan("");
                else if (xm > 0.0) worst = fmax(worst, rm / (anorm * xm));
                else if (rm > 0.0) worst = __builtin_inf();
            }
            worst = block_max(worst);
            const int open = s_flag;
            __syncthreads();
            if (!open) {
                st.resid = worst;
                break;
            }
            if (!(dmax <= 0.5 * prev) || it == REFINE_MAX_ITERS || isnan(worst)) {
                bad = 1;
                break;
            }
            prev = dmax;
            ++st.iters;
        }
        if (!bad) {
            wg_copy(n, nrhs, x, n, B, n);
            return st;
        }
    }
    // FP64 fallback
    st.fallback = 1;
    st.resid = 0.0;
    wg_copy(n, n, A, n, lu64, n);
    if (wg_lu(n, lu64, n, piv64)) {
        st.singular = 1;
        return st;
    }
    wg_lu_solve(n, lu64, n, piv64, B, n, nrhs);
    return st;
}
