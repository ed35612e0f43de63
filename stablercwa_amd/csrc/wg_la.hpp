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
#include "tuned.hpp"

#if SRCWA_LU_OUTLINE
#define LU_FN __device__ __noinline__
#else
#define LU_FN DEV
#endif

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

// Compiler-only barrier: keeps the LDS loads of a staged triangular block next to their use (per row
// and per loop trip) instead of hoisting all of them into registers ahead of the loop.
DEV void keep_loads_in_loop() { asm volatile("" ::: "memory"); }

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
// which each wave works on its own matrices).  CPRE = true (with MODE SUB or ADD): a tile's C values
// are loaded together with its first operands, so their latency is hidden behind the MFMAs (for short
// k, as in the LU and substitution updates).
template <int MODE, bool BT, bool WAVE = false, bool CPRE = false>
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
        cd cpre[2][2][4];
        if (CPRE && MODE != GEMM_SET) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int col = c0 + 16 * j + li;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int row = r0 + 16 * i + lk + 4 * q;
                        cpre[i][j][q] = (col < n && row < m) ? C[row + (long)col * ldc] : cd{0.0, 0.0};
                    }
                }
        }
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
                    const cd c = (CPRE && MODE != GEMM_SET) ? cpre[i][j][q] : (MODE == GEMM_SET ? v : *p);
                    if (MODE == GEMM_SET) *p = v;
                    else if (MODE == GEMM_SUB) *p = c - v;
                    else *p = c + v;
                }
            }
    }
    if (!WAVE) __syncthreads();
}

// C (m x nrhs) -= A (m x k) B (k x nrhs) for the short k of the blocked substitutions: MFMA tiles with
// C prefetched, or one VALU dot product per entry when there are too few right-hand sides to fill a
// 32-column tile.
DEV void wg_update(int m, int nrhs, int k, const cd* A, int lda, const cd* B, int ldb, cd* C, int ldc) {
    if (nrhs >= 8) {
        wg_gemm<GEMM_SUB, false, false, true>(m, nrhs, k, A, lda, B, ldb, C, ldc);
        return;
    }
    for (int e = threadIdx.x; e < m * nrhs; e += NT) {
        const int i = e % m, c = e / m;
        const cd* b = B + (long)c * ldb;
        cd s{0.0, 0.0};
        for (int kk = 0; kk < k; ++kk) s = s + A[i + (long)kk * lda] * b[kk];
        C[i + (long)c * ldc] = C[i + (long)c * ldc] - s;
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------------------
// Elementwise helpers (all end with a barrier).
//
// wg_map runs an elementwise pass with WG_MAP_U loads in flight per thread: ld(e) computes the new value
// (or values) of element e from loads only, st(e, v) stores it.  Each group's loads are issued before any of its
// stores, so a pass whose output may alias its input (which keeps the compiler from reordering) still
// overlaps the memory latency of WG_MAP_U elements instead of paying it once per element.  Passes are
// strictly elementwise: element e reads nothing another element of the pass writes.
constexpr int WG_MAP_U = 8;
template <class Load, class Store>
DEV void wg_map(int count, Load ld, Store st) {
    int e0 = threadIdx.x;
    for (; e0 + (WG_MAP_U - 1) * NT < count; e0 += WG_MAP_U * NT) {
        decltype(ld(0)) v[WG_MAP_U];
#pragma unroll
        for (int u = 0; u < WG_MAP_U; ++u) v[u] = ld(e0 + u * NT);
#pragma unroll
        for (int u = 0; u < WG_MAP_U; ++u) st(e0 + u * NT, v[u]);
    }
    for (int e = e0; e < count; e += NT) st(e, ld(e));
    __syncthreads();
}

DEV void wg_copy(int m, int n, const cd* src, int lds, cd* dst, int ldd) {
    wg_map(m * n, [&](int e) { return ldc(src + e % m + (long)(e / m) * lds); },
           [&](int e, cd v) { stc(dst + e % m + (long)(e / m) * ldd, v); });
}
DEV void wg_fill(long count, cd* dst, cd v) {
    for (long e = threadIdx.x; e < count; e += NT) stc(dst + e, v);
    __syncthreads();
}
DEV void wg_identity(int n, cd* dst, int ld) {
    for (int e = threadIdx.x; e < n * n; e += NT) {
        const int i = e % n, j = e / n;
        stc(dst + i + (long)j * ld, cd{i == j ? 1.0 : 0.0, 0.0});
    }
    __syncthreads();
}
DEV void wg_diag(int n, const cd* d, cd* dst, int ld) {
    wg_map(n * n, [&](int e) { const int i = e % n; return i == e / n ? ldc(d + i) : cd{0.0, 0.0}; },
           [&](int e, cd v) { stc(dst + e % n + (long)(e / n) * ld, v); });
}
// dst(i, j) = d[i] * src(i, j)      (diagonal times dense)
DEV void wg_scale_rows(int m, int n, const cd* d, const cd* src, int lds, cd* dst, int ldd) {
    wg_map(m * n, [&](int e) { const int i = e % m; return ldc(d + i) * ldc(src + i + (long)(e / m) * lds); },
           [&](int e, cd v) { stc(dst + e % m + (long)(e / m) * ldd, v); });
}
// dst(i, j) = src(i, j) * d[j]      (dense times diagonal)
DEV void wg_scale_cols(int m, int n, const cd* src, int lds, const cd* d, cd* dst, int ldd) {
    wg_map(m * n, [&](int e) { const int j = e / m; return ldc(src + e % m + (long)j * lds) * ldc(d + j); },
           [&](int e, cd v) { stc(dst + e % m + (long)(e / m) * ldd, v); });
}

// ------------------------------------------------------------------------------------------------
// LU with partial pivoting.  piv[j] = row exchanged with row j at step j.  Returns 1 when a pivot
// was exactly zero (the factorization is then only a failure marker).

// Largest value, ties to the smaller index, over one wave; every lane receives the winner.
DEV void wave_argmax(double& v, int& idx) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o, 64);
        const int oi = __shfl_xor(idx, o, 64);
        if (ov > v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    }
}

// Workgroup scratch of one panel step of the LU.
struct LuPanelShared {
    cd prow[LU_NB];              // the pivot row of the current column
    cd jrow[LU_NB];              // the row it replaces
    cd L11[LU_NB * LU_NB];       // unit-lower diagonal block of the finished panel (strict lower part)
    double rv[NWAVE];            // per-wave pivot candidates
    int ri[NWAVE];
    int porig, jorig;            // original panel rows of the two exchanged rows
    int piv[LU_NB];              // pivot row of each column (panel-relative)
    int src[LU_NB];              // panel row k (< nb) now holds what was panel row src[k]
    int pdst[LU_NB], psrc[LU_NB];// rows >= nb that changed: pdst[j] now holds what was psrc[j]
    int npairs;
    int sing;
};

// Panel factorization with the panel in registers: thread t holds panel rows t, t + NT, ..., t + (R-1) NT
// (all LU_NB columns).  Per column: pivot search (wave reduction + one barrier), exchange of the pivot
// row through LDS (one barrier), multipliers and the rank-1 update in registers.  The pivot rule, the
// multipliers (by 1/pivot when finite, else by division) and the update order are those of the
// column-at-a-time LU.  Leaves in S the pivots, L11 and the row permutation of the panel.
template <int R>
DEV void lu_panel_regs(int m, int nb, cd* Ap, int lda, LuPanelShared& S) {
    const int tid = threadIdx.x;
    cd v[R][LU_NB];
    int orig[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = tid + r * NT;
        orig[r] = row;
#pragma unroll
        for (int c = 0; c < LU_NB; ++c) {
            const cd x = ldc(Ap + (row < m ? row : 0) + (long)(c < nb ? c : 0) * lda);
            v[r][c] = (row < m && c < nb) ? x : cd{0.0, 0.0};
        }
    }
    int sing = 0;
    // the column loop stays rolled; column jj of a register row is picked by unrolled selects, so the
    // rows never need dynamic register indexing
    for (int jj = 0; jj < nb; ++jj) {
        double best = -1.0;
        int bi = m;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = tid + r * NT;
            cd vj = v[r][0];
#pragma unroll
            for (int c = 1; c < LU_NB; ++c)
                if (c == jj) vj = v[r][c];
            if (row >= jj && row < m) {
                const double a = cabs(vj);
                if (a > best) {
                    best = a;
                    bi = row;
                }
            }
        }
        wave_argmax(best, bi);
        if (lane_id() == 0) {
            S.rv[wave_id()] = best;
            S.ri[wave_id()] = bi;
        }
        __syncthreads();
        double bv = S.rv[0];
        int p = S.ri[0];
#pragma unroll
        for (int w = 1; w < NWAVE; ++w)
            if (S.rv[w] > bv || (S.rv[w] == bv && S.ri[w] < p)) {
                bv = S.rv[w];
                p = S.ri[w];
            }
        if (p >= m) p = jj;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = tid + r * NT;
            if (row == p) {
#pragma unroll
                for (int c = 0; c < LU_NB; ++c) S.prow[c] = v[r][c];
                S.porig = orig[r];
            }
            if (row == jj && p != jj) {
#pragma unroll
                for (int c = 0; c < LU_NB; ++c) S.jrow[c] = v[r][c];
                S.jorig = orig[r];
            }
        }
        if (tid == 0) S.piv[jj] = p;
        __syncthreads();
        const cd pivot = S.prow[jj];
        const bool zero = is_zero(pivot);
        sing |= zero;
        const cd inv = zero ? cd{0.0, 0.0} : cdiv(cd{1.0, 0.0}, pivot);
        const bool fin = finite(inv);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = tid + r * NT;
            if (p != jj) {
                if (row == jj) {
#pragma unroll
                    for (int c = 0; c < LU_NB; ++c) v[r][c] = S.prow[c];
                    orig[r] = S.porig;
                } else if (row == p) {
#pragma unroll
                    for (int c = 0; c < LU_NB; ++c) v[r][c] = S.jrow[c];
                    orig[r] = S.jorig;
                }
            }
            if (row > jj && row < m) {
                cd vj = v[r][0];
#pragma unroll
                for (int c = 1; c < LU_NB; ++c)
                    if (c == jj) vj = v[r][c];
                const cd l = zero ? cd{0.0, 0.0} : (fin ? vj * inv : cdiv(vj, pivot));
#pragma unroll
                for (int c = 0; c < LU_NB; ++c) {
                    if (c == jj) v[r][c] = l;
                    else if (c > jj) v[r][c] = v[r][c] - l * S.prow[c];     // zeros past nb, never stored
                }
            }
        }
    }
    // write back; L11 and the permutation to LDS
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = tid + r * NT;
        if (row < m) {
#pragma unroll
            for (int c = 0; c < LU_NB; ++c) stc(Ap + row + (long)(c < nb ? c : 0) * lda, c < nb ? v[r][c] : v[r][0]);
        }
        if (row < nb) {
#pragma unroll
            for (int c = 0; c < LU_NB; ++c) S.L11[row + c * LU_NB] = c < row ? v[r][c] : cd{0.0, 0.0};
            S.src[row] = orig[r];
        } else if (row < m && orig[r] != row) {
            const int k = atomicAdd(&S.npairs, 1);
            S.pdst[k] = row;
            S.psrc[k] = orig[r];
        }
    }
    if (tid == 0 && sing) S.sing = 1;
}

// Panels of up to 64 RL rows, factored by wave 0 alone with the panel in its registers: lane l holds
// rows l, l + 64, ... (all LU_NB columns).  Per column: pivot search by a wave reduction, the exchange of
// the two rows and the broadcast of the pivot row through LDS, multipliers and the rank-1 update in
// registers -- no workgroup barrier at all.  The column loop is unrolled, so every register index is
// static.  Same pivots, multipliers and update order as lu_panel_regs; the caller's barrier publishes
// the results to the other waves.
template <int RL>
DEV void lu_panel_wave(int m, int nb, cd* Ap, int lda, LuPanelShared& S) {
    const int lane = lane_id();
    cd v[RL][LU_NB];
    int orig[RL];
#pragma unroll
    for (int q = 0; q < RL; ++q) {
        const int row = lane + 64 * q;
        orig[q] = row;
        // clamped addresses and selects instead of guarded loads: a guard on c < nb would turn the
        // loop into one with a run-time trip count and the rows into private memory
        const int rr = row < m ? row : 0;
#pragma unroll
        for (int c = 0; c < LU_NB; ++c) {
            const cd x = ldc(Ap + rr + (long)(c < nb ? c : 0) * lda);
            v[q][c] = (row < m && c < nb) ? x : cd{0.0, 0.0};
        }
    }
    int sing = 0;
#pragma unroll
    for (int jj = 0; jj < LU_NB; ++jj) {
        if (jj < nb) {
            double best = -1.0;
            int bi = m;
#pragma unroll
            for (int q = 0; q < RL; ++q) {
                const int row = lane + 64 * q;
                if (row >= jj && row < m) {
                    const double a = cabs(v[q][jj]);
                    if (a > best) {
                        best = a;
                        bi = row;
                    }
                }
            }
            wave_argmax(best, bi);
            const int p = bi >= m ? jj : bi;
#pragma unroll
            for (int q = 0; q < RL; ++q) {
                const int row = lane + 64 * q;
                if (row == p) {
#pragma unroll
                    for (int c = 0; c < LU_NB; ++c) stc(&S.prow[c], v[q][c]);
                    S.porig = orig[q];
                }
                if (row == jj && p != jj) {
#pragma unroll
                    for (int c = 0; c < LU_NB; ++c) stc(&S.jrow[c], v[q][c]);
                    S.jorig = orig[q];
                }
            }
            if (lane == 0) S.piv[jj] = p;
            wave_sync();
            const cd pivot = ldc(&S.prow[jj]);
            const bool zero = is_zero(pivot);
            sing |= zero;
            const cd inv = zero ? cd{0.0, 0.0} : cdiv(cd{1.0, 0.0}, pivot);
            const bool fin = finite(inv);
#pragma unroll
            for (int q = 0; q < RL; ++q) {
                const int row = lane + 64 * q;
                if (p != jj) {
                    if (row == jj) {
#pragma unroll
                        for (int c = 0; c < LU_NB; ++c) v[q][c] = ldc(&S.prow[c]);
                        orig[q] = S.porig;
                    } else if (row == p) {
#pragma unroll
                        for (int c = 0; c < LU_NB; ++c) v[q][c] = ldc(&S.jrow[c]);
                        orig[q] = S.jorig;
                    }
                }
                if (row > jj && row < m) {
                    const cd l = zero ? cd{0.0, 0.0} : (fin ? v[q][jj] * inv : cdiv(v[q][jj], pivot));
                    // columns past nb hold zeros in every row (the pivot row included) and are never
                    // stored, so they are updated without a guard; the loop has a fixed trip count so it
                    // unrolls before the column loop does
#pragma unroll
                    for (int c = 0; c < LU_NB; ++c) {
                        if (c == jj) v[q][c] = l;
                        else if (c > jj) v[q][c] = v[q][c] - l * ldc(&S.prow[c]);
                    }
                }
            }
            wave_sync();     // every lane has read the exchange buffers before the next column writes them
        }
    }
#pragma unroll
    for (int q = 0; q < RL; ++q) {
        const int row = lane + 64 * q;
        if (row < m) {
            // past nb, column 0 is stored again with its own value
#pragma unroll
            for (int c = 0; c < LU_NB; ++c) stc(Ap + row + (long)(c < nb ? c : 0) * lda, c < nb ? v[q][c] : v[q][0]);
        }
        if (row < nb) {
#pragma unroll
            for (int c = 0; c < LU_NB; ++c) S.L11[row + c * LU_NB] = c < row ? v[q][c] : cd{0.0, 0.0};
            S.src[row] = orig[q];
        } else if (row < m && orig[q] != row) {
            const int k = atomicAdd(&S.npairs, 1);
            S.pdst[k] = row;
            S.psrc[k] = orig[q];
        }
    }
    if (lane == 0 && sing) S.sing = 1;
}

// Same outputs as lu_panel_regs for panels too tall for registers: one wave factors the panel in
// place in global memory, then thread 0 derives the panel's row permutation from the pivots.
DEV void lu_panel_global(int m, int nb, cd* Ap, int lda, LuPanelShared& S, int* trow, int* tsrc) {
    const int tid = threadIdx.x, lane = lane_id();
    if (wave_id() == 0) {
        int sing = 0;
        for (int jj = 0; jj < nb; ++jj) {
            cd* cj = Ap + (long)jj * lda;
            double best = -1.0;
            int bi = m;
            for (int i = jj + lane; i < m; i += 64) {
                const double a = cabs(cj[i]);
                if (a > best) {
                    best = a;
                    bi = i;
                }
            }
            wave_argmax(best, bi);
            const int p = bi >= m ? jj : bi;
            if (lane == 0) S.piv[jj] = p;
            if (p != jj)
                for (int c = lane; c < nb; c += 64) {
                    cd* col = Ap + (long)c * lda;
                    const cd t = col[jj];
                    col[jj] = col[p];
                    col[p] = t;
                }
            wave_sync();
            const cd pivot = cj[jj];
            const bool zero = is_zero(pivot);
            sing |= zero;
            const cd inv = zero ? cd{0.0, 0.0} : cdiv(cd{1.0, 0.0}, pivot);
            const bool fin = finite(inv);
            for (int i = jj + 1 + lane; i < m; i += 64) {
                const cd l = zero ? cd{0.0, 0.0} : (fin ? cj[i] * inv : cdiv(cj[i], pivot));
                cj[i] = l;
                for (int c = jj + 1; c < nb; ++c) Ap[i + (long)c * lda] = Ap[i + (long)c * lda] - l * Ap[jj + (long)c * lda];
            }
            wave_sync();
        }
        if (lane == 0 && sing) S.sing = 1;
    }
    __syncthreads();
    if (tid < LU_NB * LU_NB) {
        const int i = tid % LU_NB, c = tid / LU_NB;
        S.L11[tid] = (i < nb && c < i) ? Ap[i + (long)c * lda] : cd{0.0, 0.0};
    }
    if (tid == 0) {
        // rows touched by the panel's exchanges: trow[t] now holds what was tsrc[t]
        int cnt = nb;
        for (int k = 0; k < nb; ++k) {
            trow[k] = k;
            tsrc[k] = k;
        }
        for (int k = 0; k < nb; ++k) {
            const int p = S.piv[k];
            int b = -1;
            for (int t = 0; t < cnt; ++t)
                if (trow[t] == p) b = t;
            if (b < 0) {
                b = cnt++;
                trow[b] = p;
                tsrc[b] = p;
            }
            const int x = tsrc[k];
            tsrc[k] = tsrc[b];
            tsrc[b] = x;
        }
        for (int k = 0; k < nb; ++k) S.src[k] = tsrc[k];
        int np = 0;
        for (int t = nb; t < cnt; ++t)
            if (tsrc[t] != trow[t]) {
                S.pdst[np] = trow[t];
                S.psrc[np] = tsrc[t];
                ++np;
            }
        S.npairs = np;
    }
}

// The other columns of a full panel (nb = LU_NB), one column per thread: the panel's row permutation
// as a gather (loads of every source row first, then the stores; the list of moved rows below the
// block is padded with row 0, whose slot is rewritten last) and, right of the panel, U12 = L11^{-1} A12
// in registers.  Straight-line code: every index is static.
DEV void lu_columns_full(int n, int ks, cd* A, int lda, const LuPanelShared& S) {
    const int ke = ks + LU_NB;
    for (int c = threadIdx.x; c < n - LU_NB; c += NT) {
        keep_loads_in_loop();
        const int ci = c < ks ? c : c + LU_NB;
        cd* col = A + ks + (long)ci * lda;
        cd x[LU_NB], y[LU_NB];
#pragma unroll
        for (int k = 0; k < LU_NB; ++k) x[k] = ldc(col + S.src[k]);
#pragma unroll
        for (int j = 0; j < LU_NB; ++j) y[j] = ldc(col + S.psrc[j]);
#pragma unroll
        for (int j = 0; j < LU_NB; ++j) stc(col + S.pdst[j], y[j]);
        if (ci >= ke) {
#pragma unroll
            for (int i = 1; i < LU_NB; ++i) {
                keep_loads_in_loop();
                cd s = x[i];
#pragma unroll
                for (int k = 0; k < i; ++k) s = s - S.L11[i + k * LU_NB] * x[k];
                x[i] = s;
            }
        }
#pragma unroll
        for (int k = 0; k < LU_NB; ++k) stc(col + k, x[k]);
    }
}

// The same for a last panel narrower than LU_NB: the exchanges one at a time, then the solve.
DEV void lu_columns_partial(int n, int ks, int nb, cd* A, int lda, const LuPanelShared& S) {
    const int ke = ks + nb;
    for (int c = threadIdx.x; c < n - nb; c += NT) {
        const int ci = c < ks ? c : c + nb;
        cd* col = A + ks + (long)ci * lda;
        for (int k = 0; k < nb; ++k) {
            const int p = S.piv[k];
            if (p != k) {
                const cd t = col[k];
                col[k] = col[p];
                col[p] = t;
            }
        }
        if (ci >= ke)
            for (int i = 1; i < nb; ++i) {
                cd s = col[i];
                for (int k = 0; k < i; ++k) s = s - S.L11[i + k * LU_NB] * col[k];
                col[i] = s;
            }
    }
}

// P A = L U in place, right-looking with panels of LU_NB columns and partial pivoting (largest
// modulus, first on ties).  Per panel: the panel factorization (in one wave's registers for up to 128
// rows, in the workgroup's registers for up to 2 NT),
// then one pass over the other columns that applies the panel's row permutation and, right of the
// panel, solves U12 = L11^{-1} A12; then the trailing MFMA update.
LU_FN int wg_lu(int n, cd* A, int lda, int* piv) {
    __shared__ LuPanelShared S;
    __shared__ int s_trow[2 * LU_NB], s_tsrc[2 * LU_NB];
    const int tid = threadIdx.x;
    if (tid == 0) S.sing = 0;
    for (int ks = 0; ks < n; ks += LU_NB) {
        const int m = n - ks;
        const int nb = min(LU_NB, m);
        const int ke = ks + nb;
        cd* Ap = A + ks + (long)ks * lda;
        if (tid == 0) S.npairs = 0;
        if (tid < LU_NB) {
            S.pdst[tid] = 0;
            S.psrc[tid] = 0;
        }
        __syncthreads();
        if (m <= 128) {
            if (wave_id() == 0) {
                if (m <= 64) lu_panel_wave<1>(m, nb, Ap, lda, S);
                else lu_panel_wave<2>(m, nb, Ap, lda, S);
            }
        } else if (m <= NT) lu_panel_regs<1>(m, nb, Ap, lda, S);
        else if (m <= 2 * NT) lu_panel_regs<2>(m, nb, Ap, lda, S);
        else lu_panel_global(m, nb, Ap, lda, S, s_trow, s_tsrc);
        __syncthreads();
        if (tid < nb) piv[ks + tid] = ks + S.piv[tid];
        if (nb == LU_NB) lu_columns_full(n, ks, A, lda, S);
        else lu_columns_partial(n, ks, nb, A, lda, S);
        __syncthreads();
        if (ke < n)
            wg_gemm<GEMM_SUB, false, false, true>(n - ke, n - ke, nb, A + ke + (long)ks * lda, lda, A + ks + (long)ke * lda,
                                                  lda, A + ke + (long)ke * lda, lda);
    }
    __syncthreads();
    const int sing = S.sing;
    __syncthreads();   // every thread has read the flag before a later call clears it
    return sing;
}

// Row permutation of wg_lu's pivots applied to B (n x nrhs): the rows that move are staged through LDS in
// column blocks.
constexpr int SOLVE_STAGE_WORDS = 1024;
constexpr int SOLVE_MAX_PERM = 2048;

// Solve A X = B in place from wg_lu's factors; B is n x nrhs with leading dimension ldb.
// Blocked by LU_NB: each diagonal block is staged in LDS and solved in registers (one right-hand side
// per thread), the off-diagonal updates are MFMA products (VALU products for fewer than 8 right-hand
// sides).  The substitution order within a block is that of plain forward / backward substitution.
LU_FN void wg_lu_solve(int n, const cd* LU, int lda, const int* piv, cd* B, int ldb, int nrhs) {
    __shared__ int s_perm[SOLVE_MAX_PERM];
    __shared__ int s_moved[SOLVE_MAX_PERM];
    __shared__ int s_nmoved;
    __shared__ cd s_stage[SOLVE_STAGE_WORDS];
    __shared__ cd s_T[LU_NB * LU_NB];
    const int tid = threadIdx.x;
    // row exchanges
    if (n <= SOLVE_MAX_PERM && n <= SOLVE_STAGE_WORDS) {
        for (int i = tid; i < n; i += NT) s_perm[i] = piv[i];
        __syncthreads();
        if (tid == 0) {
            // the exchanges applied in order to the identity: cur[i] = original row now at row i; only
            // the rows that end up holding another row move
            int* cur = reinterpret_cast<int*>(s_stage);
            for (int i = 0; i < n; ++i) cur[i] = i;
            for (int k = 0; k < n; ++k) {
                const int p = s_perm[k];
                if (p != k) {
                    const int t = cur[k];
                    cur[k] = cur[p];
                    cur[p] = t;
                }
            }
            int nm = 0;
            for (int i = 0; i < n; ++i) {
                s_perm[i] = cur[i];
                if (cur[i] != i) s_moved[nm++] = i;
            }
            s_nmoved = nm;
        }
        __syncthreads();
        const int nm = s_nmoved;
        if (nm > 0) {
            const int cc = SOLVE_STAGE_WORDS / nm;
            for (int c0 = 0; c0 < nrhs; c0 += cc) {
                const int w = min(cc, nrhs - c0) * nm;
                wg_map(w, [&](int e) { return ldc(B + s_perm[s_moved[e % nm]] + (long)(c0 + e / nm) * ldb); },
                       [&](int e, cd v) { s_stage[e] = v; });
                wg_map(w, [&](int e) { return s_stage[e]; },
                       [&](int e, cd v) { stc(B + s_moved[e % nm] + (long)(c0 + e / nm) * ldb, v); });
            }
        }
    } else {
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
    }
    // forward substitution with the unit-lower factor
    for (int kb = 0; kb < n; kb += LU_NB) {
        const int ke = min(kb + LU_NB, n), nb = ke - kb;
        {
            const int i = tid % LU_NB, k = tid / LU_NB;
            s_T[tid] = (i < nb && k < i) ? LU[kb + i + (long)(kb + k) * lda] : cd{0.0, 0.0};
        }
        __syncthreads();
        if (nb == LU_NB) {
            for (int c = tid; c < nrhs; c += NT) {
                keep_loads_in_loop();
                cd* col = B + kb + (long)c * ldb;
                cd x[LU_NB];
#pragma unroll
                for (int k = 0; k < LU_NB; ++k) x[k] = ldc(col + k);
#pragma unroll
                for (int i = 1; i < LU_NB; ++i) {
                    keep_loads_in_loop();
                    cd s = x[i];
#pragma unroll
                    for (int k = 0; k < i; ++k) s = s - s_T[i + k * LU_NB] * x[k];
                    x[i] = s;
                }
#pragma unroll
                for (int k = 1; k < LU_NB; ++k) stc(col + k, x[k]);
            }
        } else {
            for (int c = tid; c < nrhs; c += NT) {
                cd* col = B + kb + (long)c * ldb;
                for (int i = 1; i < nb; ++i) {
                    cd s = col[i];
                    for (int k = 0; k < i; ++k) s = s - s_T[i + k * LU_NB] * col[k];
                    col[i] = s;
                }
            }
        }
        __syncthreads();
        if (ke < n) wg_update(n - ke, nrhs, nb, LU + ke + (long)kb * lda, lda, B + kb, ldb, B + ke, ldb);
    }
    // backward substitution with the upper factor
    const int last = ((n - 1) / LU_NB) * LU_NB;
    for (int kb = last; kb >= 0; kb -= LU_NB) {
        const int ke = min(kb + LU_NB, n), nb = ke - kb;
        {
            const int i = tid % LU_NB, k = tid / LU_NB;
            s_T[tid] = (i < nb && k < nb && k >= i) ? LU[kb + i + (long)(kb + k) * lda] : cd{0.0, 0.0};
        }
        __syncthreads();
        if (nb == LU_NB) {
            for (int c = tid; c < nrhs; c += NT) {
                keep_loads_in_loop();
                cd* col = B + kb + (long)c * ldb;
                cd x[LU_NB];
#pragma unroll
                for (int k = 0; k < LU_NB; ++k) x[k] = ldc(col + k);
#pragma unroll
                for (int i = LU_NB - 1; i >= 0; --i) {
                    keep_loads_in_loop();
                    cd s = x[i];
#pragma unroll
                    for (int k = i + 1; k < LU_NB; ++k) s = s - s_T[i + k * LU_NB] * x[k];
                    x[i] = cdiv(s, s_T[i + i * LU_NB]);
                }
#pragma unroll
                for (int k = 0; k < LU_NB; ++k) stc(col + k, x[k]);
            }
        } else {
            for (int c = tid; c < nrhs; c += NT) {
                cd* col = B + kb + (long)c * ldb;
                for (int i = nb - 1; i >= 0; --i) {
                    cd s = col[i];
                    for (int k = i + 1; k < nb; ++k) s = s - s_T[i + k * LU_NB] * col[k];
                    col[i] = cdiv(s, s_T[i + i * LU_NB]);
                }
            }
        }
        __syncthreads();
        if (kb > 0) wg_update(kb, nrhs, nb, LU + (long)kb * lda, lda, B + kb, ldb, B, ldb);
    }
}
