// Crossed-grating (2-D) RCWA on gfx950 (the reference's solve2d, rcwa2d.hpp:1042-1411).
//
//   k_setup2d     one workgroup per instance: Kx, Ky, kz of the outer media, the Rayleigh check of
//                 every homogeneous medium, and the homogeneous modes (q, 2x2 block-diagonal V)
//                                                              (rcwa2d.hpp:380-404, 1076-1130)
//   k_ops2d       one workgroup per patterned layer: Eps, Cx, Cy by the Li factorisation; the small
//                 per-segment Toeplitz inverses run one per thread  (rcwa2d_operators.hpp:349-429)
//   k_pq2d        one workgroup per patterned layer: Ei = Eps^-1, P, Q, H = P Q (2N x 2N)
//                                                              (rcwa2d.hpp:412-474)
//   k_eig2d       one workgroup per patterned layer: eigenpairs of H (CORTH + COMQR2)
//   k_modes2d     one workgroup per patterned layer: q = branch(sqrt(w)), V = (Q W) / q
//                                                              (rcwa2d.hpp:499-533)
//   k_smatrix2d   one workgroup per instance: the direct total-reflection recursion
//                 r' = A+^-1 A- with the 2x2 block-diagonal fast paths, the top block K d = 2e,
//                 the downward sweep, efficiencies and s/p amplitudes (rcwa2d.hpp:560-910, 1180-1267)
#include "batch.hpp"
#include "eig.hpp"
#include "wave_eig.hpp"
#include "rcwa_common.hpp"

// ------------------------------------------------------------------------------------------------
// flat crossed-grating stack record: Lx, Ly, nI, nII, L, then per layer d, kind, uniform: n |
// patterned: Ky, y_edges[Ky+1], then per y segment Kx, x_edges[Kx+1], n[Kx]
struct Layer2DView {
    double d;
    bool homogeneous;
    cd n;
    int K;                  // y segments
    const double* y_edges;
    const double* rows;     // first x profile record
};

DEV int64_t row_words(const double* p) {
    const int Kx = (int)p[0];
    return 1 + (Kx + 1) + 2 * Kx;
}

DEV Layer2DView layer2d_at(const double* st, int index) {
    int64_t k = 6;
    Layer2DView v{};
    for (int j = 0; j <= index; ++j) {
        v.d = st[k];
        const int kind = (int)st[k + 1];
        k += 2;
        if (kind == 0) {
            v.homogeneous = true;
            v.n = cd{st[k], st[k + 1]};
            v.K = 0;
            v.y_edges = nullptr;
            v.rows = nullptr;
            k += 2;
        } else {
            v.homogeneous = false;
            v.K = (int)st[k];
            v.y_edges = st + k + 1;
            v.rows = st + k + 2 + v.K;
            k += 1 + (v.K + 1);
            const double* r = v.rows;
            for (int y = 0; y < v.K; ++y) {
                const int64_t w = row_words(r);
                k += w;
                r += w;
            }
        }
    }
    return v;
}

// kz / k0 = sqrt(n^2 - Kx^2 - Ky^2) with Im >= 0 (rcwa2d.hpp:380-385)
DEV cd kz_over_k0(cd n, double Kx, double Ky) {
    cd kz = csqrt(n * n - cd{Kx * Kx, 0.0} - cd{Ky * Ky, 0.0});
    if (kz.y < 0.0) kz = -kz;
    return kz;
}

// q (2N) and V = Qb / q1 as a 2x2 block diagonal (rcwa2d.hpp:393-404)
DEV void homogeneous_modes2d(cd n, const cd* geo, int N, cd* q, cd* vb) {
    const cd n2 = n * n;
    for (int i = threadIdx.x; i < N; i += NT) {
        const double Kx = geo[i].x, Ky = geo[N + i].x;
        const cd kz = kz_over_k0(n, Kx, Ky);
        const cd q1 = branch_q(cd{kz.y, -kz.x});
        q[i] = q1;
        q[N + i] = q1;
        vb[i] = cdiv(cd{Kx * Ky, 0.0}, q1);
        vb[N + i] = cdiv(n2 - cd{Kx * Kx, 0.0}, q1);
        vb[2 * N + i] = cdiv(cd{Ky * Ky, 0.0} - n2, q1);
        vb[3 * N + i] = cdiv(cd{-Ky * Kx, 0.0}, q1);
    }
}

// ------------------------------------------------------------------------------------------------
#if SRCWA_MAIN_TU
__global__ __launch_bounds__(NT) void k_setup2d(const Inst2D* __restrict__ insts, const double* __restrict__ stacks,
                                                const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                double* __restrict__ geometry) {
    const Inst2D in = insts[blockIdx.x];
    const double* st = stacks + in.stack_off;
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, L = in.L, n2 = 2 * in.M2 + 1;
    const double Lx = st[0], Ly = st[1], nI = st[2];
    const cd nII{st[3], st[4]};
    const double th = in.theta_deg * (kPi / 180.0), ph = in.phi_deg * (kPi / 180.0);
    const double sin_th = sin(th), cos_ph = cos(ph), sin_ph = sin(ph);
    cd* geo = ws + lay[0];
    int zero = 0;
    for (int i = threadIdx.x; i < N; i += NT) {
        const int m = i / n2 - in.M1, nn = i % n2 - in.M2;
        const double Kx = nI * sin_th * cos_ph - m * in.wavelength / Lx;
        const double Ky = nI * sin_th * sin_ph - nn * in.wavelength / Ly;
        geo[i] = cd{Kx, 0.0};
        geo[N + i] = cd{Ky, 0.0};
        const cd kzI = kz_over_k0(cd{nI, 0.0}, Kx, Ky), kzII = kz_over_k0(nII, Kx, Ky);
        geo[2 * N + i] = kzI;
        geo[3 * N + i] = kzII;
        if (geometry) {
            double* o = geometry + GEO_WORDS * (in.out_off + i);
            o[0] = Kx;
            o[1] = Ky;
            o[2] = kzI.x;
            o[3] = kzI.y;
            o[4] = kzII.x;
            o[5] = kzII.y;
            o[6] = 0.0;
            o[7] = 0.0;
        }
        zero |= is_zero(kzI) || is_zero(kzII);
    }
    // Rayleigh anomaly in any homogeneous medium (rcwa2d.hpp:1101-1114)
    for (int j = 0; j < L; ++j) {
        const Layer2DView lv = layer2d_at(st, j);
        if (!lv.homogeneous) continue;
        for (int i = threadIdx.x; i < N; i += NT) {
            const int m = i / n2 - in.M1, nn = i % n2 - in.M2;
            const double Kx = nI * sin_th * cos_ph - m * in.wavelength / Lx;
            const double Ky = nI * sin_th * sin_ph - nn * in.wavelength / Ly;
            zero |= is_zero(kz_over_k0(lv.n, Kx, Ky));
        }
    }
    zero = block_max_int(zero);
    if (threadIdx.x == 0) geo[4 * N] = cd{(double)zero, 0.0};
    __syncthreads();
    homogeneous_modes2d(cd{nI, 0.0}, geo, N, ws + lay[x2_q(0)], ws + lay[x2_v(0)]);
    homogeneous_modes2d(nII, geo, N, ws + lay[x2_q(L + 1)], ws + lay[x2_v(L + 1)]);
    for (int j = 0; j < L; ++j) {
        if (lay[x2_slot(L, j)] >= 0) continue;
        const Layer2DView lv = layer2d_at(st, j);
        homogeneous_modes2d(lv.n, geo, N, ws + lay[x2_q(j + 1)], ws + lay[x2_v(j + 1)]);
    }
    __syncthreads();
}
#endif

// ------------------------------------------------------------------------------------------------
// Operators.
//   c[h + hmax], h = -hmax..hmax: Fourier coefficients of a piecewise-constant profile
//   (rcwa2d_operators.hpp:78-95); one thread.
DEV void thread_fourier_1d(const double* edges, const cd* values, int K, double period, int hmax, cd* c) {
    cd mean{0.0, 0.0};
    for (int j = 0; j < K; ++j) mean = mean + values[j] * (edges[j + 1] - edges[j]);
    c[hmax] = cdiv(mean, period);
    for (int h = -hmax; h <= hmax; ++h) {
        if (h == 0) continue;
        const double w = -2.0 * kPi * (double)h;
        cd sum{0.0, 0.0};
        double ps, pc;
        sincos((-2.0 * kPi * ((double)h * edges[0])) / period, &ps, &pc);
        for (int j = 0; j < K; ++j) {
            double s1, c1;
            sincos((-2.0 * kPi * ((double)h * edges[j + 1])) / period, &s1, &c1);
            sum = sum + values[j] * cd{c1 - pc, s1 - ps};
            ps = s1;
            pc = c1;
        }
        c[h + hmax] = cd{sum.y / w, -sum.x / w};
    }
}

// out(k, h + hmax) of segment k (rcwa2d_operators.hpp:103-128); out is K x (2 hmax + 1), column-major
DEV void wg_seg_fourier(const double* edges, int K, int hmax, double period, cd* out) {
    const int H = 2 * hmax + 1;
    for (int e = threadIdx.x; e < K * H; e += NT) {
        const int k = e % K, hh = e / K;
        const int h = hh - hmax;
        cd v;
        if (h == 0) {
            v = cd{(edges[k + 1] - edges[k]) / period, 0.0};
        } else {
            const double w = -2.0 * kPi * (double)h;
            double s0, c0, s1, c1;
            sincos((-2.0 * kPi * ((double)h * edges[k])) / period, &s0, &c0);
            sincos((-2.0 * kPi * ((double)h * edges[k + 1])) / period, &s1, &c1);
            const cd d{c1 - c0, s1 - s0};
            v = cd{d.y / w, -d.x / w};
        }
        out[e] = v;
    }
}

// Inverse of an n x n matrix by partial-pivot LU and n right-hand sides, one thread (the small
// per-segment Toeplitz inverses).  a is destroyed.  Returns 1 on an exact zero pivot.
DEV int thread_inverse(int n, cd* a, cd* inv, int* piv) {
    for (int j = 0; j < n; ++j) {
        int p = j;
        double best = cabs(a[j + j * n]);
        for (int i = j + 1; i < n; ++i) {
            const double v = cabs(a[i + j * n]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        piv[j] = p;
        if (p != j)
            for (int c = 0; c < n; ++c) {
                const cd t = a[j + c * n];
                a[j + c * n] = a[p + c * n];
                a[p + c * n] = t;
            }
        const cd pivot = a[j + j * n];
        if (is_zero(pivot)) return 1;
        const cd r = cdiv(cd{1.0, 0.0}, pivot);
        const bool fin = finite(r);
        for (int i = j + 1; i < n; ++i) a[i + j * n] = fin ? a[i + j * n] * r : cdiv(a[i + j * n], pivot);
        for (int c = j + 1; c < n; ++c) {
            const cd u = a[j + c * n];
            for (int i = j + 1; i < n; ++i) a[i + c * n] = a[i + c * n] - a[i + j * n] * u;
        }
    }
    for (int c = 0; c < n; ++c) {
        cd* x = inv + c * n;
        for (int i = 0; i < n; ++i) x[i] = cd{i == c ? 1.0 : 0.0, 0.0};
        for (int k = 0; k < n; ++k) {
            const int p = piv[k];
            if (p != k) {
                const cd t = x[k];
                x[k] = x[p];
                x[p] = t;
            }
        }
        for (int i = 1; i < n; ++i) {
            cd s = x[i];
            for (int k = 0; k < i; ++k) s = s - a[i + k * n] * x[k];
            x[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            cd s = x[i];
            for (int k = i + 1; k < n; ++k) s = s - a[i + k * n] * x[k];
            x[i] = cdiv(s, a[i + i * n]);
        }
    }
    return 0;
}

DEV double floor_mod(double a, double b) {
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0.0) != (m < 0.0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

// eps = n^2 of an x profile at x, wrapped by the period (rcwa2d_operators.hpp:150-160)
DEV cd eps_at(const double* row, double x, double Lx) {
    const int Kx = (int)row[0];
    const double* xe = row + 1;
    const double* v = row + 2 + Kx;
    const double xx = xe[0] + floor_mod(x - xe[0], Lx);
    int k = -1;
    for (int i = 0; i <= Kx; ++i)
        if (xe[i] <= xx) k = i;            // searchsorted(side="right") - 1
    k = min(max(k, 0), Kx - 1);
    const cd n{v[2 * k], v[2 * k + 1]};
    return n * n;
}

// host-sized scratch for the operator assembly (see ops2d_scratch_words in the runtime)
struct Ops2DScratch {
    cd *wseg, *lau, *invx, *wsegx, *invy, *work, *blocks, *coef;
    int *piv;
    const double** rows;
};

#if SRCWA_MAIN_TU
// Eps, Cx, Cy of one patterned layer (Li factorization or the Laurent rule) into EPS, CX, CY, with the
// scratch region sc (ops2d_scratch_words); returns 1 if a per-segment inverse failed or a value is not finite.
DEV int build_ops2d(const Inst2D& in, const double* st, const Layer2DView& lv, const double* cuts_rec, cd* EPS,
                    cd* CX, cd* CY, cd* sc) {
    const int M1 = in.M1, M2 = in.M2, N = in.N;
    const int n1 = 2 * M1 + 1, n2 = 2 * M2 + 1;
    const int64_t NN = (int64_t)N * N;
    const double Lx = st[0], Ly = st[1];
    const int K = lv.K;
    const int Kc = (int)cuts_rec[0];
    const double* xcuts = cuts_rec + 1;
    const int64_t n1s = (int64_t)n1 * n1, n2s = (int64_t)n2 * n2;
    cd* wseg = sc;
    cd* lau = wseg + (int64_t)K * (4 * M2 + 1);
    cd* invx = lau + K * n1s;
    cd* wsegx = invx + K * n1s;
    cd* invy = wsegx + (int64_t)Kc * (4 * M1 + 1);
    cd* work = invy + Kc * n2s;                  // one matrix per thread-task
    const int64_t worksz = max((int64_t)K * n1s, (int64_t)Kc * n2s);
    cd* blocks = work + worksz;
    const int64_t blocksz = max((int64_t)(4 * M2 + 1) * n1s, (int64_t)(4 * M1 + 1) * n2s);
    cd* coef = blocks + blocksz;                  // per thread-task 4M+1 coefficients + K values
    // x-profile record of every y segment (the host limits K to 256)
    __shared__ const double* s_rows[256];
    __shared__ int s_maxkx;
    if (threadIdx.x == 0) {
        const double* r = lv.rows;
        int mk = 1;
        for (int k = 0; k < K && k < 256; ++k) {
            s_rows[k] = r;
            mk = max(mk, (int)r[0]);
            r += row_words(r);
        }
        s_maxkx = mk;
    }
    __syncthreads();
    const int64_t coefsz = (int64_t)max(s_maxkx + 4 * M1 + 1, K + 4 * M2 + 1);
    int* piv = reinterpret_cast<int*>(coef + coefsz * max(K, Kc));
    int fail = 0;
    wg_seg_fourier(lv.y_edges, K, 2 * M2, Ly, wseg);
    // per y segment: [[eps]]_k and, for the Li rule, [[1/eps]]_k^-1 (one thread per segment)
    for (int k = threadIdx.x; k < K; k += NT) {
        const double* row = s_rows[k];
        const int Kx = (int)row[0];
        const double* xe = row + 1;
        const double* v = row + 2 + Kx;
        cd* vals = coef + (int64_t)k * coefsz;
        cd* c = vals + Kx;
        for (int i = 0; i < Kx; ++i) {
            const cd nn{v[2 * i], v[2 * i + 1]};
            vals[i] = nn * nn;
        }
        thread_fourier_1d(xe, vals, Kx, Lx, 2 * M1, c);
        cd* t = lau + k * n1s;
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n1; ++i) t[i + j * n1] = c[i - j + 2 * M1];
        if (!in.laurent) {
            for (int i = 0; i < Kx; ++i) vals[i] = cdiv(cd{1.0, 0.0}, vals[i]);
            thread_fourier_1d(xe, vals, Kx, Lx, 2 * M1, c);
            cd* w = work + k * n1s;
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n1; ++i) w[i + j * n1] = c[i - j + 2 * M1];
            fail |= thread_inverse(n1, w, invx + k * n1s, piv + k * n1);
        }
    }
    __syncthreads();
    // Eps_h[h] = sum_k wseg(k, h) [[eps]]_k, then the xfirst assembly (rcwa2d_operators.hpp:202-219)
    for (int pass = 0; pass < (in.laurent ? 1 : 2); ++pass) {
        const cd* mats = pass == 0 ? lau : invx;
        const int H = 4 * M2 + 1;
        for (int64_t e = threadIdx.x; e < (int64_t)H * n1s; e += NT) {
            const int h = (int)(e / n1s);
            const int64_t t = e % n1s;
            cd acc{0.0, 0.0};
            for (int k = 0; k < K; ++k) acc = acc + wseg[k + (int64_t)h * K] * mats[k * n1s + t];
            blocks[e] = acc;
        }
        __syncthreads();
        cd* out = pass == 0 ? EPS : CX;
        for (int64_t e = threadIdx.x; e < NN; e += NT) {
            const int row = (int)(e % N), col = (int)(e / N);
            const int mm = row / n2, nn = row % n2, mp = col / n2, np = col % n2;
            out[e] = blocks[(int64_t)(nn - np + 2 * M2) * n1s + mm + (int64_t)mp * n1];
        }
        __syncthreads();
    }
    if (in.laurent) {
        wg_copy(N, N, EPS, N, CX, N);
        wg_copy(N, N, EPS, N, CY, N);
    } else {
        // Cy: y-direction inverse rule on every x-cut segment, Laurent sum along x (rcwa2d_operators.hpp:405-422)
        wg_seg_fourier(xcuts, Kc, 2 * M1, Lx, wsegx);
        for (int j = threadIdx.x; j < Kc; j += NT) {
            const double xm = 0.5 * (xcuts[j] + xcuts[j + 1]);
            cd* vals = coef + (int64_t)j * coefsz;
            cd* c = vals + K;
            for (int k = 0; k < K; ++k) vals[k] = cdiv(cd{1.0, 0.0}, eps_at(s_rows[k], xm, Lx));
            thread_fourier_1d(lv.y_edges, vals, K, Ly, 2 * M2, c);
            cd* w = work + j * n2s;
            for (int b = 0; b < n2; ++b)
                for (int a = 0; a < n2; ++a) w[a + b * n2] = c[a - b + 2 * M2];
            fail |= thread_inverse(n2, w, invy + j * n2s, piv + j * n2);
        }
        __syncthreads();
        const int H = 4 * M1 + 1;
        for (int64_t e = threadIdx.x; e < (int64_t)H * n2s; e += NT) {
            const int h = (int)(e / n2s);
            const int64_t t = e % n2s;
            cd acc{0.0, 0.0};
            for (int j = 0; j < Kc; ++j) acc = acc + wsegx[j + (int64_t)h * Kc] * invy[j * n2s + t];
            blocks[e] = acc;
        }
        __syncthreads();
        for (int64_t e = threadIdx.x; e < NN; e += NT) {
            const int row = (int)(e % N), col = (int)(e / N);
            const int mm = row / n2, nn = row % n2, mp = col / n2, np = col % n2;
            CY[e] = blocks[(int64_t)(mm - mp + 2 * M1) * n2s + nn + (int64_t)np * n2];
        }
        __syncthreads();
    }
    int bad = 0;
    for (int64_t e = threadIdx.x; e < NN; e += NT) bad |= !finite(EPS[e]) || !finite(CX[e]) || !finite(CY[e]);
    return block_max_int(bad) | block_max_int(fail);
}

// A crossed-grating operator set (layout: crossed_ops_words): Eps, Cx, Cy, the LU factors of Eps, Ei = Eps^-1,
// the pivots, then the operator scratch.  Its failure word: bit 0 from build_ops2d, bit 1 a singular Eps.
struct CrossedOps {
    cd *EPS, *CX, *CY, *LU, *Ei, *sc;
    int* piv;
};
DEV CrossedOps crossed_ops_at(cd* base, int N) {
    const int64_t m = ops_align16((int64_t)N * N);
    CrossedOps o;
    o.EPS = base;
    o.CX = base + m;
    o.CY = base + 2 * m;
    o.LU = base + 3 * m;
    o.Ei = base + 4 * m;
    o.piv = reinterpret_cast<int*>(base + 5 * m);
    o.sc = base + 5 * m + ops_align16(N);
    return o;
}
DEV bool shares_ops2d(const LayerTask& tk, const Inst2D& in) { return tk.op >= 0 && !(in.flags & PLAN_DIAGNOSTICS); }

// The wavelength-invariant operators of one patterned layer, with the arithmetic of the per-task form
// (k_ops2d, and Ei as in k_pq2d; rcwa2d_operators.hpp:349-429, rcwa2d.hpp:412-420).
__global__ __launch_bounds__(NT) void k_ops2d_shared(const LayerOps* __restrict__ descs, const Inst2D* __restrict__ insts,
                                                     const double* __restrict__ stacks,
                                                     const int64_t* __restrict__ layouts, cd* __restrict__ ops,
                                                     int* __restrict__ ops_info) {
    const LayerOps d = descs[blockIdx.x];
    const Inst2D in = insts[d.inst];
    const double* st = stacks + in.stack_off;
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N;
    const CrossedOps o = crossed_ops_at(ops + d.off, N);
    const int bad = build_ops2d(in, st, layer2d_at(st, d.layer), stacks + lay[x2_cuts(in.L, d.layer)], o.EPS, o.CX,
                                o.CY, o.sc);
    wg_copy(N, N, o.EPS, N, o.LU, N);
    const int fail = wg_lu(N, o.LU, N, o.piv);
    wg_identity(N, o.Ei, N);
    if (!fail) wg_lu_solve(N, o.LU, N, o.piv, o.Ei, N, N);
    if (threadIdx.x == 0) ops_info[blockIdx.x] = (bad ? 1 : 0) | (fail ? 2 : 0);
}

__global__ __launch_bounds__(NT) void k_ops2d(const Inst2D* __restrict__ insts, const LayerTask* __restrict__ tasks,
                                              const double* __restrict__ stacks, const int64_t* __restrict__ layouts,
                                              cd* __restrict__ ws, int* __restrict__ task_info,
                                              const int* __restrict__ ops_info) {
    const LayerTask tk = tasks[blockIdx.x];
    const Inst2D in = insts[tk.inst];
    const double* st = stacks + in.stack_off;
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, L = in.L;
    const int64_t NN = (int64_t)N * N;
    if (threadIdx.x == 0) task_info[4 * blockIdx.x + 1] = task_info[4 * blockIdx.x + 2] = task_info[4 * blockIdx.x + 3] = 0;
    if (ws[lay[0] + 4 * N].x != 0.0) {           // Rayleigh instance: nothing to assemble
        if (threadIdx.x == 0) task_info[4 * blockIdx.x] = 0;
        return;
    }
    if (shares_ops2d(tk, in)) {                  // the layer's operators are in its shared set
        if (threadIdx.x == 0) task_info[4 * blockIdx.x] = (ops_info[tk.op] & 1) ? -1 : 0;
        return;
    }
    cd* base = ws + lay[x2_slot(L, tk.layer)];
    const int bad = build_ops2d(in, st, layer2d_at(st, tk.layer), stacks + lay[x2_cuts(L, tk.layer)], base,
                                base + NN, base + 2 * NN, base + 3 * NN);   // scratch: the BIG region
    if (threadIdx.x == 0) task_info[4 * blockIdx.x] = bad ? -1 : 0;
}
#endif

// ------------------------------------------------------------------------------------------------
struct Slots2D {
    cd *EPS, *CX, *CY, *H, *Z, *X, *lam, *ort, *tmp;
    int* piv;
    cd* D;          // with diagnostics: a copy of H (2N x 2N) for the eigen residual
};
DEV Slots2D slots2d(cd* ws, const int64_t* lay, int L, int j, int N) {
    const int64_t NN = (int64_t)N * N, nn = 4 * NN;
    cd* base = ws + lay[x2_slot(L, j)];
    Slots2D s;
    s.EPS = base;
    s.CX = base + NN;
    s.CY = base + 2 * NN;
    s.H = base + 3 * NN;
    s.Z = s.H + nn;
    s.X = s.Z + nn;
    cd* v = base + 3 * NN + lay[x2_big(L, j)];
    s.lam = v;
    s.ort = v + 2 * N;
    s.tmp = v + 4 * N;
    s.piv = reinterpret_cast<int*>(v + 6 * N);
    s.D = base + ((3 * NN + lay[x2_big(L, j)] + 7 * N + 15) & ~int64_t(15));
    return s;
}

#if SRCWA_MAIN_TU
// Ei = Eps^-1, P = [pc1 | pc2], Q12 = Cy - Kx^2, Q21 = Ky^2 - Cx, H = P Q (rcwa2d.hpp:412-474)
__global__ __launch_bounds__(NT) void k_pq2d(const Inst2D* __restrict__ insts, const LayerTask* __restrict__ tasks,
                                             const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                             int* __restrict__ task_info, TaskDiag* __restrict__ task_diag,
                                             const LayerOps* __restrict__ descs, cd* __restrict__ ops,
                                             const int* __restrict__ ops_info) {
    // one eigensolve per patterned layer, one dense solve for Ei = Eps^-1 (rcwa2d.hpp:412-474)
    if (threadIdx.x == 0) task_diag[blockIdx.x] = TaskDiag{0.0, 0.0, 1, 1};
    if (task_info[4 * blockIdx.x] != 0) return;
    const LayerTask tk = tasks[blockIdx.x];
    const Inst2D in = insts[tk.inst];
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, n = 2 * N;
    const int64_t NN = (int64_t)N * N;
    const cd* geo = ws + lay[0];
    if (geo[4 * N].x != 0.0) return;
    const Slots2D s = slots2d(ws, lay, in.L, tk.layer, N);
    // Ei into Z[NN..2NN) (Z[0..NN) holds the LU of Eps), or Ei, Cx and Cy from the layer's shared set
    const cd *Ei, *CXin = s.CX, *CYin = s.CY;
    int fail;
    if (shares_ops2d(tk, in)) {
        const CrossedOps o = crossed_ops_at(ops + descs[tk.op].off, N);
        Ei = o.Ei;
        CXin = o.CX;
        CYin = o.CY;
        fail = (ops_info[tk.op] >> 1) & 1;
    } else {
        cd* lu = s.Z;
        cd* ei = s.Z + NN;
        wg_copy(N, N, s.EPS, N, lu, N);
        fail = wg_lu(N, lu, N, s.piv);
        wg_identity(N, ei, N);
        if (!fail) wg_lu_solve(N, lu, N, s.piv, ei, N, N);
        Ei = ei;
    }
    // pc1 = [Kx Ei Ky ; Ky Ei Ky - I], pc2 = [I - Kx Ei Kx ; -Ky Ei Kx], 2N x N each (ld 2N) in X
    cd* pc1 = s.X;
    cd* pc2 = s.X + 2 * NN;
    for (int64_t e = threadIdx.x; e < NN; e += NT) {
        const int i = (int)(e % N), j = (int)(e / N);
        const double Kxi = geo[i].x, Kyi = geo[N + i].x, Kxj = geo[j].x, Kyj = geo[N + j].x;
        const cd ei = Ei[e];
        const cd kxe = ei * Kxi, kye = ei * Kyi;
        const double delta = i == j ? 1.0 : 0.0;
        pc1[i + (int64_t)j * n] = kxe * Kyj;
        pc2[i + (int64_t)j * n] = cd{delta, 0.0} - kxe * Kxj;
        pc1[N + i + (int64_t)j * n] = kye * Kyj - cd{delta, 0.0};
        pc2[N + i + (int64_t)j * n] = -(kye * Kxj);
    }
    // Q12 = Cy - Kx^2 (in CY), Q21 = Ky^2 - Cx (in CX)
    for (int64_t e = threadIdx.x; e < NN; e += NT) {
        const int i = (int)(e % N), j = (int)(e / N);
        cd q12 = CYin[e], q21 = -CXin[e];
        if (i == j) {
            q12 = q12 - cd{geo[i].x * geo[i].x, 0.0};
            q21 = q21 + cd{geo[N + i].x * geo[N + i].x, 0.0};
        }
        s.CY[e] = q12;
        s.CX[e] = q21;
    }
    __syncthreads();
    // H = [pc1 Q11 + pc2 Q21 | pc1 Q12 + pc2 Q22]
    wg_gemm<GEMM_SET, false>(n, N, N, pc2, n, s.CX, N, s.H, n);
    wg_gemm<GEMM_SET, false>(n, N, N, pc1, n, s.CY, N, s.H + (int64_t)N * n, n);
    for (int64_t e = threadIdx.x; e < 2 * NN; e += NT) {
        const int i = (int)(e % n), j = (int)(e / n);
        const double d1 = geo[j].x * geo[N + j].x, d2 = -(geo[N + j].x * geo[j].x);
        s.H[i + (int64_t)j * n] = s.H[i + (int64_t)j * n] + pc1[i + (int64_t)j * n] * d1;
        s.H[i + (int64_t)(N + j) * n] = s.H[i + (int64_t)(N + j) * n] + pc2[i + (int64_t)j * n] * d2;
    }
    __syncthreads();
    int bad = 0;
    for (int64_t e = threadIdx.x; e < 4 * NN; e += NT) bad |= !finite(s.H[e]);
    for (int64_t e = threadIdx.x; e < NN; e += NT) bad |= !finite(Ei[e]);
    bad = block_max_int(bad) | fail;
    if (in.flags & PLAN_DIAGNOSTICS) wg_copy(n, n, s.H, n, s.D, n);
    if (threadIdx.x == 0) task_info[4 * blockIdx.x] = bad ? -1 : 0;
}

// Eigen-pair residual of a patterned layer's 2N eigensolve from the copy of H (diagnostics only); Z is
// scratch (free until the modes kernel writes V there).
__global__ __launch_bounds__(NT) void k_eigres2d(const Inst2D* __restrict__ insts, const LayerTask* __restrict__ tasks,
                                                 const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                 const int* __restrict__ task_info, TaskDiag* __restrict__ task_diag) {
    const LayerTask tk = tasks[blockIdx.x];
    const Inst2D in = insts[tk.inst];
    if (!(in.flags & PLAN_DIAGNOSTICS)) return;
    const int64_t* lay = layouts + in.layout_off;
    if (ws[lay[0] + 4 * in.N].x != 0.0) return;     // Rayleigh anomaly: no eigensolve
    const Slots2D s = slots2d(ws, lay, in.L, tk.layer, in.N);
    double res = __builtin_nan("");
    if (task_info[4 * blockIdx.x] == 0 && task_info[4 * blockIdx.x + 1] == 0)
        res = wg_eig_residual(2 * in.N, s.D, s.H, s.lam, s.Z);
    if (threadIdx.x == 0) task_diag[blockIdx.x].eig_res = res;
}
#endif


template <int QC>
__global__ EIG_KERNEL_BOUNDS_OF(SRCWA_EIG_OCC_CROSSED) void k_eig2d(const Inst2D* __restrict__ insts, const LayerTask* __restrict__ tasks,
                                              const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                              int* __restrict__ task_info, int max_iter, const int* __restrict__ order) {
    const int t = order ? order[blockIdx.x] : static_cast<int>(blockIdx.x);   // largest matrices first
    if (task_info[4 * t] != 0) return;
    const LayerTask tk = tasks[t];
    const Inst2D in = insts[tk.inst];
    const int64_t* lay = layouts + in.layout_off;
    if (ws[lay[0] + 4 * in.N].x != 0.0) return;
    if (2 * in.N <= WAVE_EIG_MAX_N) return;        // the wave eigensolver's matrix
    const Slots2D s = slots2d(ws, lay, in.L, tk.layer, in.N);
    const int info = wg_eig<QC>(2 * in.N, s.H, s.lam, s.Z, s.X, EigScratch{s.ort, s.tmp}, max_iter, kEps);
    if (threadIdx.x == 0) task_info[4 * t + 1] = info;
}

// One wave per layer eigenproblem of order 2N <= WAVE_EIG_MAX_N, four per workgroup.
template <int QC>
__global__ __launch_bounds__(NT) void k_eig2d_wave(int n_tasks, const Inst2D* __restrict__ insts,
                                                   const LayerTask* __restrict__ tasks,
                                                   const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                   int* __restrict__ task_info, int max_iter) {
    __shared__ double s_rot[NWAVE][3 * (64 * QC + 1)];
    const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * NWAVE + wave_id());   // wave-uniform
    if (t >= n_tasks || task_info[4 * t] != 0) return;
    const LayerTask tk = tasks[t];
    const Inst2D in = insts[tk.inst];
    const int64_t* lay = layouts + in.layout_off;
    if (ws[lay[0] + 4 * in.N].x != 0.0) return;
    if (2 * in.N > WAVE_EIG_MAX_N) return;         // the workgroup eigensolver's matrix
    const Slots2D s = slots2d(ws, lay, in.L, tk.layer, in.N);
    const int info = wave_eig<QC>(2 * in.N, s.H, s.lam, s.Z, s.X, EigScratch{s.ort, s.tmp}, s_rot[wave_id()], max_iter,
                                  kEps);
    if (lane_id() == 0) task_info[4 * t + 1] = info;
}

// QcLaunch::what: 0 = k_eig2d (workgroup per matrix), 1 = k_eig2d_wave (wave per matrix)
SRCWA_QC_DISPATCH(crossed_qc, QcLaunch)
#ifdef SRCWA_QC_ONLY
void SRCWA_QC_FN(crossed_qc)(const QcLaunch& a) {
    constexpr int QC = SRCWA_QC_ONLY;
    const Inst2D* insts = static_cast<const Inst2D*>(a.insts);
    if (a.what == 0)
        k_eig2d<QC><<<a.grid, NT, 0, a.stream>>>(insts, a.tasks, a.layouts, a.ws, a.task_info, a.max_iter, a.order);
    else if constexpr (QC <= 4)
        k_eig2d_wave<QC><<<a.grid, NT, 0, a.stream>>>(a.n_items, insts, a.tasks, a.layouts, a.ws, a.task_info,
                                                      a.max_iter);
}
#endif

#if SRCWA_MAIN_TU
// q = branch(sqrt(w)), V = (Q W) / q into Z (rcwa2d.hpp:499-533)
__global__ __launch_bounds__(NT) void k_modes2d(const Inst2D* __restrict__ insts, const LayerTask* __restrict__ tasks,
                                                const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                int* __restrict__ task_info) {
    if (task_info[4 * blockIdx.x] != 0 || task_info[4 * blockIdx.x + 1] != 0) return;
    const LayerTask tk = tasks[blockIdx.x];
    const Inst2D in = insts[tk.inst];
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, n = 2 * N;
    const int64_t NN = (int64_t)N * N;
    const cd* geo = ws + lay[0];
    if (geo[4 * N].x != 0.0) return;
    const Slots2D s = slots2d(ws, lay, in.L, tk.layer, N);
    cd* q = ws + lay[x2_q(tk.layer + 1)];
    for (int j = threadIdx.x; j < n; j += NT) q[j] = branch_q(csqrt(s.lam[j]));
    __syncthreads();
    const cd* W = s.H;
    cd* V = s.Z;
    // top = Q12 Wb, bot = Q21 Wt (rows of W: Wt = W[0:N, :], Wb = W[N:2N, :])
    wg_gemm<GEMM_SET, false>(N, n, N, s.CY, N, W + N, n, V, n);
    wg_gemm<GEMM_SET, false>(N, n, N, s.CX, N, W, n, V + N, n);
    int bad = 0;
    for (int64_t e = threadIdx.x; e < 4 * NN; e += NT) {
        const int i = (int)(e % n), j = (int)(e / n);
        const int m = i < N ? i : i - N;
        const double Kx = geo[m].x, Ky = geo[N + m].x;
        cd v = V[e];
        if (i < N) v = v + W[i + (int64_t)j * n] * (Kx * Ky);
        else v = v + W[i + (int64_t)j * n] * (-(Ky * Kx));
        v = cdiv(v, q[j]);
        V[e] = v;
        bad |= !finite(v) || !finite(W[e]);
    }
    bad = block_max_int(bad);
    if (threadIdx.x == 0) task_info[4 * blockIdx.x + 2] = bad ? -1 : 0;
}
#endif

// ------------------------------------------------------------------------------------------------
// 2 x 2 block-diagonal operators (a11, a12, a21, a22 of length Nh) and the recursion state.
DEV void b2_mul(int Nh, const cd* A, const cd* B, cd* C) {
    for (int i = threadIdx.x; i < Nh; i += NT) {
        const cd a11 = A[i], a12 = A[Nh + i], a21 = A[2 * Nh + i], a22 = A[3 * Nh + i];
        const cd b11 = B[i], b12 = B[Nh + i], b21 = B[2 * Nh + i], b22 = B[3 * Nh + i];
        C[i] = a11 * b11 + a12 * b21;
        C[Nh + i] = a11 * b12 + a12 * b22;
        C[2 * Nh + i] = a21 * b11 + a22 * b21;
        C[3 * Nh + i] = a21 * b12 + a22 * b22;
    }
    __syncthreads();
}
DEV void b2_inv(int Nh, const cd* A, cd* C) {
    for (int i = threadIdx.x; i < Nh; i += NT) {
        const cd a11 = A[i], a12 = A[Nh + i], a21 = A[2 * Nh + i], a22 = A[3 * Nh + i];
        const cd det = a11 * a22 - a12 * a21;
        C[i] = cdiv(a22, det);
        C[Nh + i] = cdiv(-a12, det);
        C[2 * Nh + i] = cdiv(-a21, det);
        C[3 * Nh + i] = cdiv(a11, det);
    }
    __syncthreads();
}
// C = s1 (I + t A) with t = +-1 applied element-wise: mode 0: I + A, 1: I - A, 2: A + I(scaled s), general helper
DEV void b2_eye_plus(int Nh, const cd* A, double sign, cd* C) {
    for (int i = threadIdx.x; i < Nh; i += NT) {
        C[i] = cd{1.0, 0.0} + sign * A[i];
        C[Nh + i] = sign * A[Nh + i];
        C[2 * Nh + i] = sign * A[2 * Nh + i];
        C[3 * Nh + i] = cd{1.0, 0.0} + sign * A[3 * Nh + i];
    }
    __syncthreads();
}
// y = A x (x, y of length 2 Nh), for nrhs columns of leading dimension 2 Nh
DEV void b2_apply(int Nh, const cd* A, const cd* x, cd* y) {
    for (int i = threadIdx.x; i < Nh; i += NT) {
        const cd x1 = x[i], x2 = x[Nh + i];
        y[i] = A[i] * x1 + A[Nh + i] * x2;
        y[Nh + i] = A[2 * Nh + i] * x1 + A[3 * Nh + i] * x2;
    }
    __syncthreads();
}
// Out = bdense(A) M (rows mixed pairwise), M is 2Nh x cols with leading dimension 2Nh
DEV void b2_rows(int Nh, const cd* A, const cd* M, int cols, cd* Out) {
    const int n = 2 * Nh;
    for (int64_t e = threadIdx.x; e < (int64_t)Nh * cols; e += NT) {
        const int i = (int)(e % Nh), c = (int)(e / Nh);
        const cd x1 = M[i + (int64_t)c * n], x2 = M[Nh + i + (int64_t)c * n];
        Out[i + (int64_t)c * n] = A[i] * x1 + A[Nh + i] * x2;
        Out[Nh + i + (int64_t)c * n] = A[2 * Nh + i] * x1 + A[3 * Nh + i] * x2;
    }
    __syncthreads();
}
// Out = M bdense(B) (columns mixed pairwise), M is 2Nh x 2Nh
DEV void b2_rmul(int Nh, const cd* M, const cd* B, cd* Out) {
    const int n = 2 * Nh;
    for (int64_t e = threadIdx.x; e < (int64_t)n * Nh; e += NT) {
        const int i = (int)(e % n), j = (int)(e / n);
        const cd x1 = M[i + (int64_t)j * n], x2 = M[i + (int64_t)(Nh + j) * n];
        Out[i + (int64_t)j * n] = x1 * B[j] + x2 * B[2 * Nh + j];
        Out[i + (int64_t)(Nh + j) * n] = x1 * B[Nh + j] + x2 * B[3 * Nh + j];
    }
    __syncthreads();
}
// M += bdense(A)
DEV void b2_add_to(int Nh, cd* M, const cd* A) {
    const int n = 2 * Nh;
    for (int i = threadIdx.x; i < Nh; i += NT) {
        M[i + (int64_t)i * n] = M[i + (int64_t)i * n] + A[i];
        M[i + (int64_t)(Nh + i) * n] = M[i + (int64_t)(Nh + i) * n] + A[Nh + i];
        M[Nh + i + (int64_t)i * n] = M[Nh + i + (int64_t)i * n] + A[2 * Nh + i];
        M[Nh + i + (int64_t)(Nh + i) * n] = M[Nh + i + (int64_t)(Nh + i) * n] + A[3 * Nh + i];
    }
    __syncthreads();
}

// interface between two homogeneous media (rcwa2d.hpp:576-589): tdd, rdu, rud, tuu (4 Nh each)
DEV void interface_bd(SmatCtx& c, int Nh, const cd* up_vb, const cd* low_vb, cd* tdd, cd* rdu, cd* rud, cd* tuu) {
    const int64_t b4 = 4 * (int64_t)Nh;
    cd* inv_low = c.bump.tmp(b4);
    cd* nn = c.bump.tmp(b4);
    cd* tp = c.bump.tmp(b4);
    cd* tm = c.bump.tmp(b4);
    cd* t1 = c.bump.tmp(b4);
    c.check();
    if (c.fail) return;
    b2_inv(Nh, low_vb, inv_low);
    b2_mul(Nh, inv_low, up_vb, nn);
    for (int i = threadIdx.x; i < Nh; i += NT)
        for (int q = 0; q < 4; ++q) {
            const cd m = (q == 0 || q == 3) ? cd{1.0, 0.0} : cd{0.0, 0.0};
            tp[q * Nh + i] = 0.5 * (m + nn[q * Nh + i]);
            tm[q * Nh + i] = 0.5 * (m - nn[q * Nh + i]);
        }
    __syncthreads();
    b2_inv(Nh, tp, tuu);                 // tpi
    b2_mul(Nh, tm, tuu, rdu);            // tm tpi
    b2_mul(Nh, rdu, tm, t1);             // (tm tpi) tm
    for (int64_t e = threadIdx.x; e < b4; e += NT) tdd[e] = tp[e] - t1[e];
    __syncthreads();
    b2_mul(Nh, tuu, tm, rud);
    for (int64_t e = threadIdx.x; e < b4; e += NT) rud[e] = -rud[e];
    __syncthreads();
}

enum { R2_ZERO = 0, R2_BD = 1, R2_DENSE = 2 };
enum { STEP_BD = 1, STEP_IFACE_BD = 3, STEP_IFACE = 4 };

struct Region2D {
    bool homog;
    cd* q;
    cd* vb;
    cd* W;
    cd* V;
};

#if SRCWA_IN_PART(1)
__global__ SMAT_KERNEL_BOUNDS_OF(SRCWA_SMAT_OCC_CROSSED) void k_smatrix2d(const Inst2D* __restrict__ insts, const double* __restrict__ stacks,
                                                  const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                  const int* __restrict__ task_info, Out1D out) {
    __shared__ int s_kind[SMAT_MAX_BLOCKS];
    __shared__ cd* s_a[SMAT_MAX_BLOCKS];
    __shared__ cd* s_b[SMAT_MAX_BLOCKS];
    __shared__ cd* s_c[SMAT_MAX_BLOCKS];
    const Inst2D in = insts[blockIdx.x];
    const double* st = stacks + in.stack_off;
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, L = in.L, n = 2 * N;
    const int n2c = 2 * in.M2 + 1;
    const int64_t nn = (int64_t)n * n, b4 = 4 * (int64_t)N;
    const double k0 = 2.0 * kPi / in.wavelength;
    const double nI = st[2];
    const cd nII{st[3], st[4]};
    const double th = in.theta_deg * (kPi / 180.0), ph = in.phi_deg * (kPi / 180.0);
    const double cos_th = cos(th), cos_ph = cos(ph), sin_ph = sin(ph);
    const cd* geo = ws + lay[0];
    // per instance R = [Rs, Rp, Ex_r, Ey_r], T = [Ts, Tp, Ex_t, Ey_t] (N complex each)
    cd* Rs = reinterpret_cast<cd*>(out.R) + 4 * in.out_off;
    cd* Rp = Rs + N;
    cd* Ts = reinterpret_cast<cd*>(out.T) + 4 * in.out_off;
    cd* Tp = Ts + N;
    double* DEr = out.DEr + in.out_off;
    double* DEt = out.DEt + in.out_off;

    int status = ST_OK, eig_info = 0;
    double max_res = 0.0;
    int n_solve = 0, n_eig = 0;
    if (geo[4 * N].x != 0.0) status = ST_RAYLEIGH;      // found before any layer work: nothing counted
    if (status == ST_OK) {
        for (int t = 0; t < in.n_tasks; ++t) {
            const int* ti = task_info + 4 * (in.first_task + t);
            if (ti[0] < 0 || ti[2] < 0) status = ST_NUMERICAL;
            if (ti[1] != 0 && eig_info == 0) eig_info = ti[1];
            const TaskDiag d = out.task_diag[in.first_task + t];
            n_solve += d.n_solve;
            n_eig += d.n_eig;
            if (in.flags & PLAN_DIAGNOSTICS) max_res = max_keep_nan(max_res, d.eig_res);
        }
        if (eig_info) status = ST_NUMERICAL;
    }
    SmatCtx c;
    c.N = n;
    c.bump = Bump(ws, in.ws_lo, in.ws_hi);
    c.fail = 0;
    c.mixed = in.flags & PLAN_MIXED;
    auto region = [&](int r) -> Region2D {
        Region2D g;
        g.q = ws + lay[x2_q(r)];
        if (r == 0 || r == L + 1 || lay[x2_slot(L, r - 1)] < 0) {
            g.homog = true;
            g.vb = ws + lay[x2_v(r)];
            g.W = g.V = nullptr;
        } else {
            const Slots2D s = slots2d(ws, lay, L, r - 1, N);
            g.homog = false;
            g.vb = nullptr;
            g.W = s.H;
            g.V = s.Z;
        }
        return g;
    };
    cd* rslot[2] = {c.bump.keep(nn), c.bump.keep(nn)};
    int rs = 0;
    int rkind = R2_ZERO;
    cd* r = nullptr;
    c.check();
    if (status == ST_OK && !c.fail) {
        for (int b = 2 * L; b >= 1 && !c.fail; --b) {
            const int64_t mark = c.bump.lo;
            const int j = b / 2;
            int kind = 0;
            cd *sa = nullptr, *sb = nullptr, *sc = nullptr;
            cd* rn = nullptr;
            int rnk = R2_ZERO;
            if (b % 2 == 0) {
                const Region2D up = region(j), low = region(j + 1);
                if (up.homog && low.homog) {
                    if (rkind != R2_DENSE) {
                        cd* tdd = c.bump.tmp(b4);
                        cd* rdu = c.bump.tmp(b4);
                        cd* rud = c.bump.tmp(b4);
                        cd* tuu = c.bump.tmp(b4);
                        cd* t1 = c.bump.tmp(b4);
                        cd* t2 = c.bump.tmp(b4);
                        c.check();
                        if (c.fail) break;
                        interface_bd(c, N, up.vb, low.vb, tdd, rdu, rud, tuu);
                        // gt = (I - rdu rr)^-1 tdd; r' = rud + tuu rr gt
                        cd* gt = c.bump.keep(b4);
                        cd* rnew = c.bump.tmp(b4);
                        c.check();
                        if (c.fail) break;
                        if (rkind == R2_ZERO) {
                            for (int64_t e = threadIdx.x; e < b4; e += NT) t1[e] = cd{0.0, 0.0};
                            __syncthreads();
                        } else {
                            b2_mul(N, rdu, r, t1);
                        }
                        b2_eye_plus(N, t1, -1.0, t2);
                        b2_inv(N, t2, t1);
                        b2_mul(N, t1, tdd, gt);
                        if (rkind == R2_ZERO) {
                            for (int64_t e = threadIdx.x; e < b4; e += NT) rnew[e] = rud[e];
                            __syncthreads();
                        } else {
                            b2_mul(N, tuu, r, t1);
                            b2_mul(N, t1, gt, t2);
                            for (int64_t e = threadIdx.x; e < b4; e += NT) rnew[e] = rud[e] + t2[e];
                            __syncthreads();
                        }
                        kind = STEP_BD;
                        sa = gt;
                        rn = rnew;
                        rnk = R2_BD;
                    } else {
                        // direct_bd_over_dense (rcwa2d.hpp:790-808)
                        cd* inv_low = c.bump.tmp(b4);
                        cd* nnb = c.bump.keep(b4);
                        cd* tmb = c.bump.keep(b4);
                        cd* np1 = c.bump.tmp(b4);
                        cd* ap = c.bump.tmp(nn);
                        cd* am = c.bump.keep(nn);
                        c.check();
                        if (c.fail) break;
                        b2_inv(N, low.vb, inv_low);
                        b2_mul(N, inv_low, up.vb, nnb);
                        for (int i = threadIdx.x; i < N; i += NT)
                            for (int q = 0; q < 4; ++q) {
                                const cd m = (q == 0 || q == 3) ? cd{1.0, 0.0} : cd{0.0, 0.0};
                                np1[q * N + i] = nnb[q * N + i] + m;
                                inv_low[q * N + i] = nnb[q * N + i] - m;       // nm1
                                tmb[q * N + i] = inv_low[q * N + i] * (-0.5);   // Tm = (I - Nn)/2
                            }
                        __syncthreads();
                        b2_rmul(N, r, inv_low, ap);        // r (Nn - I)
                        b2_add_to(N, ap, np1);             // A+ = (Nn + I) + r (Nn - I)
                        b2_rmul(N, r, np1, am);            // r (Nn + I)
                        b2_add_to(N, am, inv_low);         // A- = (Nn - I) + r (Nn + I)
                        dense_solve_inplace(c, ap, am, n);
                        if (c.fail) break;
                        kind = STEP_IFACE_BD;
                        sa = tmb;
                        sb = nnb;
                        sc = am;
                        rn = am;
                        rnk = R2_DENSE;
                    }
                } else {
                    // interface_matching + direct_dense (rcwa2d.hpp:560-608, 753-788)
                    cd* mm = c.bump.keep(nn);
                    cd* nnm = c.bump.keep(nn);
                    c.check();
                    if (c.fail) break;
                    if (low.homog) {
                        cd* inv_low = c.bump.tmp(b4);
                        c.check();
                        if (c.fail) break;
                        wg_copy(n, n, up.W, n, mm, n);
                        b2_inv(N, low.vb, inv_low);
                        b2_rows(N, inv_low, up.V, n, nnm);
                    } else if (up.homog) {
                        wg_identity(n, mm, n);
                        dense_solve(c, low.W, mm, n);
                        if (c.fail) break;
                        for (int64_t e = threadIdx.x; e < nn; e += NT) nnm[e] = cd{0.0, 0.0};
                        __syncthreads();
                        b2_add_to(N, nnm, up.vb);
                        dense_solve(c, low.V, nnm, n);
                    } else {
                        wg_copy(n, n, up.W, n, mm, n);
                        dense_solve(c, low.W, mm, n);
                        if (c.fail) break;
                        wg_copy(n, n, up.V, n, nnm, n);
                        dense_solve(c, low.V, nnm, n);
                    }
                    if (c.fail) break;
                    cd* pn = c.bump.tmp(nn);
                    cd* qm = c.bump.tmp(nn);
                    c.check();
                    if (c.fail) break;
                    if (rkind == R2_ZERO) {
                        wg_copy(n, n, nnm, n, pn, n);
                        wg_copy(n, n, mm, n, qm, n);
                    } else if (rkind == R2_BD) {
                        cd* ip = c.bump.tmp(b4);
                        c.check();
                        if (c.fail) break;
                        b2_eye_plus(N, r, 1.0, ip);
                        b2_rows(N, ip, nnm, n, pn);
                        b2_eye_plus(N, r, -1.0, ip);
                        b2_rows(N, ip, mm, n, qm);
                    } else {
                        wg_gemm<GEMM_SET, false>(n, n, n, r, n, nnm, n, pn, n);
                        for (int64_t e = threadIdx.x; e < nn; e += NT) pn[e] = pn[e] + nnm[e];
                        __syncthreads();
                        wg_gemm<GEMM_SET, false>(n, n, n, r, n, mm, n, qm, n);
                        for (int64_t e = threadIdx.x; e < nn; e += NT) qm[e] = mm[e] - qm[e];
                        __syncthreads();
                    }
                    cd* rp = c.bump.keep(nn);
                    c.check();
                    if (c.fail) break;
                    for (int64_t e = threadIdx.x; e < nn; e += NT) {
                        const cd a = pn[e], q = qm[e];
                        qm[e] = a + q;         // A+
                        rp[e] = a - q;         // A-
                    }
                    __syncthreads();
                    dense_solve_inplace(c, qm, rp, n);
                    if (c.fail) break;
                    for (int64_t e = threadIdx.x; e < nn; e += NT) mm[e] = (mm[e] - nnm[e]) * 0.5;
                    __syncthreads();
                    kind = STEP_IFACE;
                    sa = mm;
                    sb = nnm;
                    sc = rp;
                    rn = rp;
                    rnk = R2_DENSE;
                }
            } else {
                // propagation (rcwa2d.hpp:615-632): tdd = tuu = diag(exp(-k0 q d)), rdu = rud = 0
                const Region2D g = region(j + 1);
                const Layer2DView lv = layer2d_at(st, j);
                cd* tdd = c.bump.keep(b4);
                c.check();
                if (c.fail) break;
                int bad = 0;
                for (int i = threadIdx.x; i < N; i += NT) {
                    tdd[i] = cexp((-k0 * g.q[i]) * lv.d);
                    tdd[N + i] = cd{0.0, 0.0};
                    tdd[2 * N + i] = cd{0.0, 0.0};
                    tdd[3 * N + i] = cexp((-k0 * g.q[N + i]) * lv.d);
                }
                __syncthreads();
                (void)bad;
                kind = STEP_BD;
                sa = tdd;
                if (rkind != R2_DENSE) {
                    cd* rnew = c.bump.tmp(b4);
                    c.check();
                    if (c.fail) break;
                    if (rkind == R2_ZERO) {
                        for (int64_t e = threadIdx.x; e < b4; e += NT) rnew[e] = cd{0.0, 0.0};
                        __syncthreads();
                    } else {
                        cd* t1 = c.bump.tmp(b4);
                        c.check();
                        if (c.fail) break;
                        b2_mul(N, tdd, r, t1);         // tuu r
                        b2_mul(N, t1, tdd, rnew);      // (tuu r) gt
                    }
                    rn = rnew;
                    rnk = R2_BD;
                } else {
                    cd* t1 = c.bump.tmp(nn);
                    cd* rnew = c.bump.tmp(nn);
                    c.check();
                    if (c.fail) break;
                    b2_rows(N, tdd, r, n, t1);
                    b2_rmul(N, t1, tdd, rnew);
                    rn = rnew;
                    rnk = R2_DENSE;
                }
            }
            if (threadIdx.x == 0) {
                s_kind[b] = kind;
                s_a[b] = sa;
                s_b[b] = sb;
                s_c[b] = sc;
            }
            // persist r'
            cd* dst = rslot[rs ^ 1];
            const int64_t cnt = rnk == R2_DENSE ? nn : b4;
            if (rn != dst) {
                for (int64_t e = threadIdx.x; e < cnt; e += NT) dst[e] = rn[e];
                __syncthreads();
            }
            r = dst;
            rkind = rnk;
            rs ^= 1;
            c.bump.lo = mark;
            __syncthreads();
        }
        if (c.fail) status = c.fail == 2 ? ST_WORKSPACE : ST_NUMERICAL;
    }
    if (status == ST_OK) {
        // incident vector (rcwa2d.hpp:1171-1174)
        const int i0 = in.M1 * n2c + in.M2;
        const cd Es{in.jones[0], in.jones[1]}, Ep{in.jones[2], in.jones[3]};
        cd* inc = c.bump.tmp(n);
        cd* down = c.bump.tmp(n);
        cd* Rv = c.bump.tmp(n);
        cd* tv = c.bump.tmp(n);
        cd* tv2 = c.bump.tmp(n);
        c.check();
        if (!c.fail) {
            for (int i = threadIdx.x; i < n; i += NT) inc[i] = cd{0.0, 0.0};
            __syncthreads();
            if (threadIdx.x == 0) {
                inc[i0] = -Es * sin_ph + Ep * (cos_th * cos_ph);
                inc[N + i0] = Es * cos_ph + Ep * (cos_th * sin_ph);
            }
            __syncthreads();
            const Region2D up = region(0), low = region(1);
            if (low.homog) {
                // both homogeneous (rcwa2d.hpp:830-843)
                cd* tdd = c.bump.tmp(b4);
                cd* rdu = c.bump.tmp(b4);
                cd* rud = c.bump.tmp(b4);
                cd* tuu = c.bump.tmp(b4);
                cd* t1 = c.bump.tmp(b4);
                cd* t2 = c.bump.tmp(b4);
                c.check();
                if (!c.fail) {
                    interface_bd(c, N, up.vb, low.vb, tdd, rdu, rud, tuu);
                    if (rkind != R2_DENSE) {
                        if (rkind == R2_ZERO) {
                            for (int64_t e = threadIdx.x; e < b4; e += NT) t1[e] = cd{0.0, 0.0};
                            __syncthreads();
                        } else {
                            b2_mul(N, rdu, r, t1);
                        }
                        b2_eye_plus(N, t1, -1.0, t2);
                        b2_inv(N, t2, t1);          // g
                        b2_mul(N, t1, tdd, t2);     // g tdd
                        b2_apply(N, t2, inc, down);
                        if (rkind == R2_ZERO) {
                            for (int i = threadIdx.x; i < n; i += NT) tv[i] = cd{0.0, 0.0};
                            __syncthreads();
                        } else {
                            b2_apply(N, r, down, tv);
                        }
                    } else {
                        cd* mat = c.bump.tmp(nn);
                        c.check();
                        if (!c.fail) {
                            b2_rows(N, rdu, r, n, mat);
                            for (int64_t e = threadIdx.x; e < nn; e += NT) {
                                const int i = (int)(e % n), jj = (int)(e / n);
                                cd v = -mat[e];
                                if (i == jj) v = v + cd{1.0, 0.0};
                                mat[e] = v;
                            }
                            __syncthreads();
                            b2_apply(N, tdd, inc, down);
                            dense_solve_inplace(c, mat, down, 1);
                            if (!c.fail) op_apply_cols(c, DOp{OP_DENSE, r}, down, tv, 1);
                        }
                    }
                    if (!c.fail) {
                        b2_apply(N, tuu, tv, tv2);
                        b2_apply(N, rud, inc, Rv);
                        for (int i = threadIdx.x; i < n; i += NT) Rv[i] = Rv[i] + tv2[i];
                        __syncthreads();
                    }
                }
            } else {
                // direct top block K d = 2e, R = W (I + r) d - e (rcwa2d.hpp:846-872)
                cd* wp = c.bump.tmp(nn);
                cd* vq = c.bump.tmp(nn);
                cd* kmat = c.bump.tmp(nn);
                cd* inv_up = c.bump.tmp(b4);
                c.check();
                if (!c.fail) {
                    if (rkind == R2_ZERO) {
                        wg_copy(n, n, low.W, n, wp, n);
                        wg_copy(n, n, low.V, n, vq, n);
                    } else if (rkind == R2_BD) {
                        cd* ip = inv_up;
                        b2_eye_plus(N, r, 1.0, ip);
                        b2_rmul(N, low.W, ip, wp);
                        b2_eye_plus(N, r, -1.0, ip);
                        b2_rmul(N, low.V, ip, vq);
                    } else {
                        wg_gemm<GEMM_SET, false>(n, n, n, low.W, n, r, n, wp, n);
                        for (int64_t e = threadIdx.x; e < nn; e += NT) wp[e] = wp[e] + low.W[e];
                        __syncthreads();
                        wg_gemm<GEMM_SET, false>(n, n, n, low.V, n, r, n, vq, n);
                        for (int64_t e = threadIdx.x; e < nn; e += NT) vq[e] = low.V[e] - vq[e];
                        __syncthreads();
                    }
                    b2_inv(N, up.vb, inv_up);
                    b2_rows(N, inv_up, vq, n, kmat);
                    for (int64_t e = threadIdx.x; e < nn; e += NT) kmat[e] = kmat[e] + wp[e];
                    for (int i = threadIdx.x; i < n; i += NT) down[i] = inc[i] * 2.0;
                    __syncthreads();
                    dense_solve_inplace(c, kmat, down, 1);
                    if (!c.fail) {
                        op_apply_cols(c, DOp{OP_DENSE, wp}, down, Rv, 1);
                        for (int i = threadIdx.x; i < n; i += NT) Rv[i] = Rv[i] - inc[i];
                        __syncthreads();
                    }
                }
            }
            // downward sweep (rcwa2d.hpp:647-664)
            for (int b = 1; b <= 2 * L && !c.fail; ++b) {
                const int kind = s_kind[b];
                if (kind == STEP_BD) {
                    b2_apply(N, s_a[b], down, tv);
                } else {
                    op_apply_cols(c, DOp{OP_DENSE, s_c[b]}, down, tv2, 1);
                    for (int i = threadIdx.x; i < n; i += NT) tv2[i] = tv2[i] + down[i];
                    __syncthreads();
                    if (kind == STEP_IFACE_BD) {
                        b2_apply(N, s_a[b], tv2, tv);
                        b2_apply(N, s_b[b], down, tv2);
                    } else {
                        op_apply_cols(c, DOp{OP_DENSE, s_a[b]}, tv2, tv, 1);
                        op_apply_cols(c, DOp{OP_DENSE, s_b[b]}, down, tv2, 1);
                    }
                    for (int i = threadIdx.x; i < n; i += NT) tv[i] = tv[i] + tv2[i];
                    __syncthreads();
                }
                cd* sw = down;
                down = tv;
                tv = sw;
            }
            if (c.fail) status = c.fail == 2 ? ST_WORKSPACE : ST_NUMERICAL;
            if (status == ST_OK) {
                // efficiencies and s/p amplitudes (rcwa2d.hpp:1232-1267)
                int bad = 0;
                for (int i = threadIdx.x; i < N; i += NT) {
                    const double Kx = geo[i].x, Ky = geo[N + i].x;
                    const double kt = hypot(Kx, Ky);
                    for (int side = 0; side < 2; ++side) {
                        const cd* F = side == 0 ? Rv : down;
                        const cd kz = side == 0 ? geo[2 * N + i] : geo[3 * N + i];
                        const double sign = side == 0 ? -1.0 : 1.0;
                        const cd nmed = side == 0 ? cd{nI, 0.0} : nII;
                        const cd Ex = F[i], Ey = F[N + i];
                        if (!finite(Ex) || !finite(Ey)) bad = 1;
                        double de = 0.0;
                        if (kz.x > 1e-12) {
                            const cd Ez = cdiv((-sign) * (Kx * Ex + Ky * Ey), kz);
                            const double E2 = norm2(Ex) + norm2(Ey) + norm2(Ez);
                            de = kz.x * E2 / (nI * cos_th);
                        }
                        const cd cth = cabs(kz) > 1e-12 ? cdiv(kz, nmed) : cd{1.0, 0.0};
                        cd sv, pv;
                        if (kt > 1e-12) {
                            sv = cdiv((-Ky) * Ex + Kx * Ey, kt);
                            pv = cdiv(sign * (Kx * Ex + Ky * Ey), cth * kt);
                        } else {
                            sv = (-sin_ph) * Ex + cos_ph * Ey;
                            pv = cdiv(sign * (cos_ph * Ex + sin_ph * Ey), cth);
                        }
                        if (!isfinite(de) || !finite(sv) || !finite(pv)) bad = 1;
                        if (side == 0) {
                            DEr[i] = de;
                            Rs[i] = sv;
                            Rp[i] = pv;
                            Rs[2 * N + i] = Ex;
                            Rs[3 * N + i] = Ey;
                        } else {
                            DEt[i] = de;
                            Ts[i] = sv;
                            Tp[i] = pv;
                            Ts[2 * N + i] = Ex;
                            Ts[3 * N + i] = Ey;
                        }
                    }
                }
                if (block_max_int(bad)) status = ST_NUMERICAL;
            }
        } else {
            status = ST_WORKSPACE;
        }
    }
    __syncthreads();
    if (status != ST_OK) {
        const double nan = __builtin_nan("");
        for (int i = threadIdx.x; i < N; i += NT) {
            for (int q = 0; q < 4; ++q) Rs[q * N + i] = Ts[q * N + i] = cd{nan, nan};
            DEr[i] = DEt[i] = nan;
        }
    }
    if (threadIdx.x == 0) {
        out.status[blockIdx.x] = status;
        out.eig_info[blockIdx.x] = eig_info;
        // Result2D reports no solve backward error (rcwa2d_types.hpp:112-130)
        double* dg = out.diag + (int64_t)DIAG_WORDS * blockIdx.x;
        dg[0] = max_res;
        dg[1] = 0.0;
        dg[2] = n_eig;
        dg[3] = n_solve + c.n_solve;
        dg[4] = 0.0;
        c.refine_record(dg);
    }
}
void smatrix2d_launch(int grid, hipStream_t s, const Inst2D* insts, const double* stacks, const int64_t* layouts, cd* ws, const int* task_info, Out1D out) {
    k_smatrix2d<<<grid, NT, 0, s>>>(insts, stacks, layouts, ws, task_info, out);
}
#endif

// ------------------------------------------------------------------------------------------------
// the S-matrix kernel is compiled in its own unit
void smatrix2d_launch(int grid, hipStream_t s, const Inst2D* insts, const double* stacks, const int64_t* layouts, cd* ws, const int* task_info, Out1D out);

#if SRCWA_MAIN_TU
void launch_crossed(hipStream_t s, int n_inst, int n_tasks, const Inst2D* insts, const LayerTask* tasks,
                    const double* stacks, const int64_t* layouts, cd* ws, int* task_info, Out1D out, int max_iter, int min_n, int max_n,
                    int part, int flags, hipEvent_t* ev, const int* eig_order, const OpsArgs& ops) {
    const bool diag = flags & PLAN_DIAGNOSTICS;
    // part 1: shared operators (first chunk), setup, assembly, eigensolves (events ev[0..2]); part 2:
    // modes, S-matrix, efficiencies (events ev[3], ev[4]); the host may run the parts of different chunks
    // on two streams
    if (part & 1) {
        if (ev) hipEventRecord(ev[0], s);
        if (ops.compute && ops.n > 0)
            k_ops2d_shared<<<ops.n, NT, 0, s>>>(ops.descs, static_cast<const Inst2D*>(ops.insts), stacks, layouts,
                                                ops.ops, ops.info);
        k_setup2d<<<n_inst, NT, 0, s>>>(insts, stacks, layouts, ws, out.geometry);
        if (n_tasks > 0) {
            k_ops2d<<<n_tasks, NT, 0, s>>>(insts, tasks, stacks, layouts, ws, task_info, ops.info);
            k_pq2d<<<n_tasks, NT, 0, s>>>(insts, tasks, layouts, ws, task_info, out.task_diag, ops.descs, ops.ops,
                                          ops.info);
        }
        if (ev) hipEventRecord(ev[1], s);
        // each matrix goes to the wave or the workgroup eigensolver by its own order
        QcLaunch a{};
        a.stream = s;
        a.n_items = n_tasks;
        a.insts = insts;
        a.tasks = tasks;
        a.layouts = layouts;
        a.ws = ws;
        a.task_info = task_info;
        a.max_iter = max_iter;
        a.order = eig_order;
        if (n_tasks > 0 && 2 * min_n <= WAVE_EIG_MAX_N) {
            a.what = 1;
            a.grid = (n_tasks + NWAVE - 1) / NWAVE;
            crossed_qc(wave_columns_per_lane(WAVE_EIG_MAX_N), a);
        }
        if (n_tasks > 0 && 2 * max_n > WAVE_EIG_MAX_N) {
            a.what = 0;
            a.grid = n_tasks;
            crossed_qc(eig_columns_per_thread(2 * max_n), a);
        }
        if (n_tasks > 0 && diag) k_eigres2d<<<n_tasks, NT, 0, s>>>(insts, tasks, layouts, ws, task_info, out.task_diag);
        if (ev) hipEventRecord(ev[2], s);
    }
    if (part & 2) {
        if (ev) hipEventRecord(ev[3], s);
        if (n_tasks > 0) k_modes2d<<<n_tasks, NT, 0, s>>>(insts, tasks, layouts, ws, task_info);
        smatrix2d_launch(n_inst, s, insts, stacks, layouts, ws, task_info, out);
        if (ev) hipEventRecord(ev[4], s);
    }
}
#endif
