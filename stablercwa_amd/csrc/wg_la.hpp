// Workgroup-level dense complex-FP64 linear algebra for gfx950.
//
// Every routine here is executed by one whole 256-thread workgroup (4 wave64s) on one matrix that
// lives in device memory; all threads must call it with identical arguments (uniform control flow),
// and it returns after a workgroup barrier, so its results are visible to every thread.  Batching
// happens one level up: a launch puts one workgroup on each instance (or each layer of an instance).
//
//   wg_gemm      C (=, -=, +=) A * B        MFMA v_mfma_f64_16x16x4_f64 tiles, 32x32 per wave
//   wg_lu        P A = L U, partial pivoting by largest modulus (first maximum wins), 16-column panels
//   wg_lu_solve  A X = B from wg_lu's factors (row swaps, blocked unit-lower and upper substitution)
//
// The LU follows the reference's lu.hpp (right-looking partial-pivot LU with column panels,
// lu.hpp:267-396; solve lu.hpp:233-256, 427-443): pivot = first row of largest |a_ij| (std::abs),
// multipliers by 1/pivot (lu.hpp:336-343), an exact zero pivot marks the factorization singular.
#pragma once

#include "cplx.hpp"

constexpr int NT = 256;          // threads per workgroup for every wg_* routine
constexpr int NWAVE = NT / 64;
constexpr int LU_NB = 16;        // LU / substitution panel width

typedef double d4 __attribute__((ext_vector_type(4)));

DEV int lane_id() { return threadIdx.x & 63; }
DEV int wave_id() { return threadIdx.x >> 6; }

DEV double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
DEV double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Orders the memory operations of the lanes of one wave (wave-scope release/acquire).
DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

DEV double read_lane(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

DEV int wave_max_int(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Sum over the workgroup; every thread receives the result.
DEV double block_sum(double v) {
    __shared__ double red[NWAVE];
    v = wave_sum(v);
    __syncthreads();
    if (lane_id() == 0) red[wave_id()] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < NWAVE; ++w) r += red[w];
    return r;
}
DEV double block_max(double v) {
    __shared__ double red[NWAVE];
    v = wave_max(v);
    __syncthreads();
    if (lane_id() == 0) red[wave_id()] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < NWAVE; ++w) r = fmax(r, red[w]);
    return r;
}
DEV int block_max_int(int v) {
    __shared__ int red[NWAVE];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if (lane_id() == 0) red[wave_id()] = v;
    __syncthreads();
    int r = red[0];
#pragma unroll
    for (int w = 1; w < NWAVE; ++w) r = max(r, red[w]);
    return r;
}
// Largest value, ties to the smaller index; every thread receives the winning index.
DEV int block_argmax(double v, int idx) {
    __shared__ double rv[NWAVE];
    __shared__ int ri[NWAVE];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o, 64);
        const int oi = __shfl_xor(idx, o, 64);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
    __syncthreads();
    if (lane_id() == 0) { rv[wave_id()] = v; ri[wave_id()] = idx; }
    __syncthreads();
    double bv = rv[0];
    int bi = ri[0];
#pragma unroll
    for (int w = 1; w < NWAVE; ++w)
        if (rv[w] > bv || (rv[w] == bv && ri[w] < bi)) { bv = rv[w]; bi = ri[w]; }
    return bi;
}

// ------------------------------------------------------------------------------------------------
// GEMM.  Each wave owns 32x32 output tiles (2x2 blocks of 16x16).  A complex product is four real
// MFMAs per block: Cre += Ar*Br, Cre += (-Ai)*Bi, Cim += Ar*Bi, Cim += Ai*Br.
// v_mfma_f64_16x16x4_f64 operand layout (per lane l): A[l%16][l/16], B[l/16][l%16];
// result element r of lane l is D[l/16 + 4*r][l%16] (checked on gfx950 against numpy).
enum { GEMM_SET = 0, GEMM_SUB = 1, GEMM_ADD = 2 };

// WAVE = true: the calling wave computes every tile alone, with no workgroup barrier (for code in
// which each wave works on its own matrices).
template <int MODE, bool BT, bool WAVE = false>
DEV void wg_gemm(int m, int n, int k, const cd* __restrict__ A, int lda, const cd* __restrict__ B, int ldb,
                 cd* C, int ldc) {
    const int lane = lane_id();
    const int li = lane & 15, lk = lane >> 4;
    const int tm = (m + 31) >> 5, tn = (n + 31) >> 5, tiles = tm * tn;
    for (int t = WAVE ? 0 : wave_id(); t < tiles; t += WAVE ? 1 : NWAVE) {
        const int r0 = (t % tm) << 5, c0 = (t / tm) << 5;
        d4 re[2][2], im[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                re[i][j] = d4{0.0, 0.0, 0.0, 0.0};
                im[i][j] = d4{0.0, 0.0, 0.0, 0.0};
            }
        const int ra[2] = {r0 + li, r0 + 16 + li};
        const int cb[2] = {c0 + li, c0 + 16 + li};
        // operands of k-step k0 (lane: row / column li, k index k0 + lk)
        auto fetch = [&](int k0, cd* a, cd* b) {
            const int kk = k0 + lk;
            const bool kin = kk < k;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a[i] = (kin && ra[i] < m) ? A[ra[i] + (long)kk * lda] : cd{0.0, 0.0};
                if (BT) b[i] = (kin && cb[i] < n) ? B[cb[i] + (long)kk * ldb] : cd{0.0, 0.0};
                else b[i] = (kin && cb[i] < n) ? B[kk + (long)cb[i] * ldb] : cd{0.0, 0.0};
            }
        };
        // the operands of the next two k-steps are in flight while the MFMAs of this one run
        cd a0[2], b0[2], a1[2], b1[2];
        fetch(0, a0, b0);
        fetch(4, a1, b1);
        for (int k0 = 0; k0 < k; k0 += 4) {
            cd a[2] = {a0[0], a0[1]}, b[2] = {b0[0], b0[1]};
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
                    re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].x, re[i][j], 0, 0, 0);
                    re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[i].y, b[j].y, re[i][j], 0, 0, 0);
                    im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].y, im[i][j], 0, 0, 0);
                    im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].x, im[i][j], 0, 0, 0);
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
                    const int row = r0 + 16 * i + lk + 4 * q;
                    if (row >= m) continue;
                    cd* p = C + row + (long)col * ldc;
                    const cd v{re[i][j][q], im[i][j][q]};
                    if (MODE == GEMM_SET) *p = v;
                    else if (MODE == GEMM_SUB) *p = *p - v;
                    else *p = *p + v;
                }
            }
    }
    if (!WAVE) __syncthreads();
}

// ------------------------------------------------------------------------------------------------
// Elementwise helpers (all end with a barrier).

DEV void wg_copy(int m, int n, const cd* src, int lds, cd* dst, int ldd) {
    for (int e = threadIdx.x; e < m * n; e += NT) {
        const int i = e % m, j = e / m;
        dst[i + (long)j * ldd] = src[i + (long)j * lds];
    }
    __syncthreads();
}
DEV void wg_fill(long count, cd* dst, cd v) {
    for (long e = threadIdx.x; e < count; e += NT) dst[e] = v;
    __syncthreads();
}
DEV void wg_identity(int n, cd* dst, int ld) {
    for (int e = threadIdx.x; e < n * n; e += NT) {
        const int i = e % n, j = e / n;
        dst[i + (long)j * ld] = cd{i == j ? 1.0 : 0.0, 0.0};
    }
    __syncthreads();
}
DEV void wg_diag(int n, const cd* d, cd* dst, int ld) {
    for (int e = threadIdx.x; e < n * n; e += NT) {
        const int i = e % n, j = e / n;
        dst[i + (long)j * ld] = i == j ? d[i] : cd{0.0, 0.0};
    }
    __syncthreads();
}
// dst(i, j) = d[i] * src(i, j)      (diagonal times dense)
DEV void wg_scale_rows(int m, int n, const cd* d, const cd* src, int lds, cd* dst, int ldd) {
    for (int e = threadIdx.x; e < m * n; e += NT) {
        const int i = e % m, j = e / m;
        dst[i + (long)j * ldd] = d[i] * src[i + (long)j * lds];
    }
    __syncthreads();
}
// dst(i, j) = src(i, j) * d[j]      (dense times diagonal)
DEV void wg_scale_cols(int m, int n, const cd* src, int lds, const cd* d, cd* dst, int ldd) {
    for (int e = threadIdx.x; e < m * n; e += NT) {
        const int i = e % m, j = e / m;
        dst[i + (long)j * ldd] = src[i + (long)j * lds] * d[j];
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------------------
// LU with partial pivoting.  piv[j] = row exchanged with row j at step j.  Returns 1 when a pivot
// was exactly zero (the factorization is then only a failure marker).
DEV int wg_lu(int n, cd* A, int lda, int* piv) {
    const int tid = threadIdx.x;
    int singular = 0;
    for (int ks = 0; ks < n; ks += LU_NB) {
        const int ke = min(ks + LU_NB, n);
        for (int j = ks; j < ke; ++j) {
            cd* cj = A + (long)j * lda;
            double best = -1.0;
            int bi = n;
            for (int i = j + tid; i < n; i += NT) {
                const double v = cabs(cj[i]);
                if (v > best) { best = v; bi = i; }
            }
            int p = block_argmax(best, bi);
            if (p >= n) p = j;
            if (tid == 0) piv[j] = p;
            if (p != j)
                for (int c = ks + tid; c < ke; c += NT) {
                    cd* col = A + (long)c * lda;
                    const cd t = col[j];
                    col[j] = col[p];
                    col[p] = t;
                }
            __syncthreads();
            const cd pivot = cj[j];
            if (is_zero(pivot)) {
                singular = 1;
                for (int i = j + 1 + tid; i < n; i += NT) cj[i] = cd{0.0, 0.0};
            } else {
                const cd inv = cdiv(cd{1.0, 0.0}, pivot);
                if (finite(inv)) {
                    for (int i = j + 1 + tid; i < n; i += NT) cj[i] = cj[i] * inv;
                } else {
                    for (int i = j + 1 + tid; i < n; i += NT) cj[i] = cdiv(cj[i], pivot);
                }
            }
            __syncthreads();
            const int rows = n - j - 1, cols = ke - j - 1;
            for (int e = tid; e < rows * cols; e += NT) {
                const int i = j + 1 + e % rows, c = j + 1 + e / rows;
                cd* col = A + (long)c * lda;
                col[i] = col[i] - cj[i] * col[j];
            }
            __syncthreads();
        }
        // the panel's row exchanges on the columns outside it
        for (int c = tid; c < n - (ke - ks); c += NT) {
            const int col_index = c < ks ? c : c + (ke - ks);
            cd* col = A + (long)col_index * lda;
            for (int k = ks; k < ke; ++k) {
                const int p = piv[k];
                if (p != k) {
                    const cd t = col[k];
                    col[k] = col[p];
                    col[p] = t;
                }
            }
        }
        __syncthreads();
        if (ke < n) {
            // U12 = L11^{-1} A12 (unit lower, one column per thread)
            for (int c = ke + tid; c < n; c += NT) {
                cd* col = A + (long)c * lda;
                for (int i = ks + 1; i < ke; ++i) {
                    cd s = col[i];
                    for (int k = ks; k < i; ++k) s = s - A[i + (long)k * lda] * col[k];
                    col[i] = s;
                }
            }
            __syncthreads();
            wg_gemm<GEMM_SUB, false>(n - ke, n - ke, ke - ks, A + ke + (long)ks * lda, lda, A + ks + (long)ke * lda, lda,
                                     A + ke + (long)ke * lda, lda);
        }
    }
    return singular;
}

// Solve A X = B in place from wg_lu's factors; B is n x nrhs with leading dimension ldb.
DEV void wg_lu_solve(int n, const cd* LU, int lda, const int* piv, cd* B, int ldb, int nrhs) {
    const int tid = threadIdx.x;
    for (int c = tid; c < nrhs; c += NT) {
        cd* col = B + (long)c * ldb;
        for (int k = 0; k < n; ++k) {
            const int p = piv[k];
            if (p != k) {
                const cd t = col[k];
                col[k] = col[p];
                col[p] = t;
            }
        }
    }
    __syncthreads();
    for (int kb = 0; kb < n; kb += LU_NB) {
        const int ke = min(kb + LU_NB, n);
        for (int c = tid; c < nrhs; c += NT) {
            cd* col = B + (long)c * ldb;
            for (int i = kb + 1; i < ke; ++i) {
                cd s = col[i];
                for (int k = kb; k < i; ++k) s = s - LU[i + (long)k * lda] * col[k];
                col[i] = s;
            }
        }
        __syncthreads();
        if (ke < n) wg_gemm<GEMM_SUB, false>(n - ke, nrhs, ke - kb, LU + ke + (long)kb * lda, lda, B + kb, ldb, B + ke, ldb);
    }
    const int last = ((n - 1) / LU_NB) * LU_NB;
    for (int kb = last; kb >= 0; kb -= LU_NB) {
        const int ke = min(kb + LU_NB, n);
        for (int c = tid; c < nrhs; c += NT) {
            cd* col = B + (long)c * ldb;
            for (int i = ke - 1; i >= kb; --i) {
                cd s = col[i];
                for (int k = i + 1; k < ke; ++k) s = s - LU[i + (long)k * lda] * col[k];
                col[i] = cdiv(s, LU[i + (long)i * lda]);
            }
        }
        __syncthreads();
        if (kb > 0) wg_gemm<GEMM_SUB, false>(kb, nrhs, ke - kb, LU + (long)kb * lda, lda, B + kb, ldb, B, ldb);
    }
}
