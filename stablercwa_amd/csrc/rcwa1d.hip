// Planar (TE/TM) RCWA on gfx950: batched kernels over instances of one launch plan.
//
//   k_setup1d      one workgroup per instance: kx_m and the analytic modes of the incidence medium,
//                  exit medium and uniform films           (rcwa1d.hpp:410-415, 522-544)
//   k_ops1d        one workgroup per shared operator set: [[eps]] (TE); the LU factors of [[eps]],
//                  [[1/eps]] and its LU factors (TM) -- the layer's wavelength-invariant operators, once
//                  for every task of the plan that has the same layer, period, M and polarization
//   k_assemble1d   one workgroup per grating layer: analytic Fourier coefficients, Toeplitz matrix,
//                  layer matrix A (TE: Kx^2 - [[eps]]; TM: [[1/eps]]^-1 (Kx [[eps]]^-1 Kx - I))
//                                                          (rcwa1d.hpp:366-407, 432-468)
//   k_eig1d        one workgroup per grating layer: eigenpairs of A (CORTH + COMQR2, eig.hpp)
//   k_modes1d      one workgroup per grating layer: q = branch(sqrt(lambda)), V = W q or [[1/eps]] W q
//                                                          (rcwa1d.hpp:949-961)
//   k_smatrix1d    one workgroup per instance: the whole S-matrix recursion with the reference's
//                  zero / diagonal / dense operator algebra, the top block, the downward sweep and
//                  the efficiencies                        (rcwa1d.hpp:549-808, 963-1055)
#include "batch.hpp"
#include "eig.hpp"
#include "wave_eig.hpp"
#include "rcwa_common.hpp"

// ------------------------------------------------------------------------------------------------
#if SRCWA_MAIN_TU
__global__ __launch_bounds__(NT) void k_setup1d(const Inst1D* __restrict__ insts, const double* __restrict__ stacks,
                                                const int64_t* __restrict__ layouts, cd* __restrict__ ws) {
    const Inst1D in = insts[blockIdx.x];
    const double* st = stacks + in.stack_off;
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, M = in.M, L = in.L;
    const double period = st[0];
    const cd nI = cd{st[1], st[2]}, nII = cd{st[3], st[4]};
    cd* kx = ws + lay[LAYOUT_KX];
    const double sn = nI.x * sin(in.theta_deg * kPi / 180.0);
    for (int j = threadIdx.x; j < N; j += NT) kx[j] = cd{sn - (j - M) * in.wavelength / period, 0.0};
    __syncthreads();
    analytic_modes_1d(nI, kx, N, in.pol, false, ws + lay[layout_q(0)], ws + lay[layout_p(0)]);
    analytic_modes_1d(nII, kx, N, in.pol, false, ws + lay[layout_q(L + 1)], ws + lay[layout_p(L + 1)]);
    LayerCursor cur(st);
    for (int j = 0; j < L; ++j) {
        const LayerView lv = cur.next();
        if (lv.homogeneous)
            analytic_modes_1d(lv.n, kx, N, in.pol, true, ws + lay[layout_q(j + 1)], ws + lay[layout_p(j + 1)]);
    }
    __syncthreads();
}
#endif

// ------------------------------------------------------------------------------------------------
#if SRCWA_MAIN_TU
DEV bool shares_ops(const LayerTask& tk, const Inst1D& in) { return tk.op >= 0 && !(in.flags & PLAN_DIAGNOSTICS); }

// A planar operator set (layout: planar_ops_words).  E: [[eps]] (TE) or its LU factors (TM).
struct PlanarOps {
    cd *E, *Ti, *LUi, *coef;
    int *pive, *pivi;
    cd* F;                                     // FFT layers: sample scratch
};
DEV PlanarOps planar_ops_at(cd* base, int N, int pol) {
    const int64_t m = ops_align16((int64_t)N * N);
    PlanarOps o{};
    o.E = base;
    cd* p = base + m;
    if (pol != 0) {
        o.Ti = p;
        o.LUi = p + m;
        o.pive = reinterpret_cast<int*>(p + 2 * m);
        o.pivi = reinterpret_cast<int*>(p + 2 * m + ops_align16(N));
        p += 2 * m + 2 * ops_align16(N);
    }
    o.coef = p;
    o.F = p + ops_align16(2 * (int64_t)N);
    return o;
}

// The wavelength-invariant operators of one grating layer, with the arithmetic of the per-task form in
// k_assemble1d (rcwa1d.hpp:432-468): the same coefficients, Toeplitz matrices and LU factorizations.
__global__ __launch_bounds__(NT) void k_ops1d(const LayerOps* __restrict__ descs, const Inst1D* __restrict__ insts,
                                              const double* __restrict__ stacks, cd* __restrict__ ops,
                                              int* __restrict__ ops_info) {
    const LayerOps d = descs[blockIdx.x];
    const Inst1D in = insts[d.inst];
    const double* st = stacks + in.stack_off;
    const int N = in.N, M = in.M;
    LayerCursor cur(st);
    LayerView lv = cur.next();
    for (int j = 0; j < d.layer; ++j) lv = cur.next();
    const PlanarOps o = planar_ops_at(ops + d.off, N, in.pol);
    int fail = layer_coefficients(lv, st[0], M, false, o.coef, o.F);
    toeplitz(o.coef, N, o.E);
    if (in.pol != 0) {
        fail |= layer_coefficients(lv, st[0], M, true, o.coef, o.F);
        toeplitz(o.coef, N, o.Ti);
        fail |= wg_lu(N, o.E, N, o.pive);
        wg_copy(N, N, o.Ti, N, o.LUi, N);
        fail |= wg_lu(N, o.LUi, N, o.pivi);
    }
    if (threadIdx.x == 0) ops_info[blockIdx.x] = fail;
}

__global__ __launch_bounds__(NT) void k_assemble1d(const Inst1D* __restrict__ insts, const LayerTask* __restrict__ tasks,
                                                   const double* __restrict__ stacks, const int64_t* __restrict__ layouts,
                                                   cd* __restrict__ ws, int* __restrict__ task_info,
                                                   TaskDiag* __restrict__ task_diag, const LayerOps* __restrict__ descs,
                                                   cd* __restrict__ ops, const int* __restrict__ ops_info) {
    const LayerTask tk = tasks[blockIdx.x];
    const Inst1D in = insts[tk.inst];
    const double* st = stacks + in.stack_off;
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, M = in.M;
    // a task with a shared operator set reads its layer's operators from it (the host gives none with
    // diagnostics, whose backward errors need the per-task matrices)
    const bool shared = shares_ops(tk, in);
    const PlanarOps o = shared ? planar_ops_at(ops + descs[tk.op].off, N, in.pol) : PlanarOps{};
    const double period = st[0];
    LayerCursor cur(st);
    LayerView lv = cur.next();
    for (int j = 0; j < tk.layer; ++j) lv = cur.next();
    const cd* kx = ws + lay[LAYOUT_KX];
    cd* base = ws + lay[layout_layer(in.L, tk.layer)];
    const int64_t NN = (int64_t)N * N;
    cd *S0 = base, *S1 = base + NN, *S2 = base + 2 * NN, *S3 = base + 3 * NN;
    cd* coef = base + 4 * NN;                 // lambda + ort hold the 4M+1 coefficients
    int* piv = reinterpret_cast<int*>(base + 4 * NN + 3 * N);
    // with diagnostics: D0 keeps the layer matrix for the eigen residual, D1 is solve scratch
    const bool diag = in.flags & PLAN_DIAGNOSTICS;
    cd* D0 = base + slot_words(N);
    cd* D1 = D0 + NN;
    cd* F = D0 + (diag ? 2 * NN : 0);          // FFT layers: 2 nfft sample scratch
    double berr = 0.0;
    int fail = shared ? ops_info[tk.op] : 0;
    if (in.pol == 0) {
        const cd* E = S0;
        if (shared) {
            E = o.E;
        } else {
            fail |= layer_coefficients(lv, period, M, false, coef, F);
            toeplitz(coef, N, S0);
        }
        // A = Kx^2 - [[eps]]
        wg_map((int)NN,
               [&](int e) {
                   const int i = e % N;
                   cd v = -ldc(E + e);
                   if (i == e / N) v = v + cd{kx[i].x * kx[i].x, 0.0};
                   return v;
               },
               [&](int e, cd v) { stc(S0 + e, v); });
    } else {
        // S1 = [[eps]], S3 = [[1/eps]] (kept for V), S2 = diag(kx); with a shared set its factors and
        // [[1/eps]] are read from there
        const cd *LUe = S1, *LUi = S0;
        const int *pive = piv, *pivi = piv;
        if (shared) {
            LUe = o.E;
            LUi = o.LUi;
            pive = o.pive;
            pivi = o.pivi;
        } else {
            fail |= layer_coefficients(lv, period, M, false, coef, F);
            toeplitz(coef, N, S1);
            fail |= layer_coefficients(lv, period, M, true, coef, F);
            toeplitz(coef, N, S3);
        }
        wg_diag(N, kx, S2, N);
        if (!shared) fail |= wg_lu(N, S1, N, piv);
        wg_lu_solve(N, LUe, N, pive, S2, N, N);
        if (diag) {
            // [[eps]] X = diag(kx): rebuild [[eps]] (S1 holds its factors) and the right-hand side
            fail |= layer_coefficients(lv, period, M, false, coef, F);
            toeplitz(coef, N, D0);
            wg_diag(N, kx, D1, N);
            berr = max_keep_nan(berr, wg_backward_error(N, D0, S2, D1, N, S0));
        }
        // S2 = Kx S2 - I
        wg_map((int)NN,
               [&](int e) {
                   const int i = e % N;
                   cd v = kx[i].x * ldc(S2 + e);
                   if (i == e / N) v = v - cd{1.0, 0.0};
                   return v;
               },
               [&](int e, cd v) { stc(S2 + e, v); });
        if (!shared) {
            wg_copy(N, N, S3, N, S0, N);
            if (diag) wg_copy(N, N, S2, N, D1, N);
            fail |= wg_lu(N, S0, N, piv);
        }
        wg_lu_solve(N, LUi, N, pivi, S2, N, N);
        if (diag) berr = max_keep_nan(berr, wg_backward_error(N, S3, S2, D1, N, D0));
        wg_copy(N, N, S2, N, S0, N);
    }
    if (diag) wg_copy(N, N, S0, N, D0, N);
    int bad = 0;
    for (int64_t e = threadIdx.x; e < NN; e += NT) bad |= !finite(S0[e]);
    bad = block_max_int(bad);
    if (threadIdx.x == 0) {
        task_info[blockIdx.x] = (fail || bad) ? -1 : 0;
        // rcwa1d.hpp:432-468: TM assembly makes two dense solves
        task_diag[blockIdx.x] = TaskDiag{0.0, berr, in.pol == 0 ? 0 : 2, 1};
    }
}

// Eigen-pair residual of a grating layer's eigensolve from the layer matrix kept in D0 (diagnostics
// only); S1 is scratch (the eigensolver's Z, free until the modes kernel writes V there).
__global__ __launch_bounds__(NT) void k_eigres1d(const Inst1D* __restrict__ insts, const LayerTask* __restrict__ tasks,
                                                 const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                 const int* __restrict__ task_info, TaskDiag* __restrict__ task_diag) {
    const LayerTask tk = tasks[blockIdx.x];
    const Inst1D in = insts[tk.inst];
    if (!(in.flags & PLAN_DIAGNOSTICS)) return;
    const int N = in.N;
    const int64_t NN = (int64_t)N * N;
    cd* base = ws + layouts[in.layout_off + layout_layer(in.L, tk.layer)];
    double res = __builtin_nan("");            // no eigenvectors: not converged, or no layer matrix
    if (task_info[blockIdx.x] == 0) res = wg_eig_residual(N, base + slot_words(N), base, base + 4 * NN, base + NN);
    if (threadIdx.x == 0) task_diag[blockIdx.x].eig_res = res;
}
#endif


// ------------------------------------------------------------------------------------------------
template <int QC>
__global__ EIG_KERNEL_BOUNDS_OF(SRCWA_EIG_OCC_PLANAR) void k_eig1d(const Inst1D* __restrict__ insts, const LayerTask* __restrict__ tasks,
                                              const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                              int* __restrict__ task_info, int max_iter) {
    if (task_info[blockIdx.x] != 0) return;        // assembly failed: nothing to decompose
    const LayerTask tk = tasks[blockIdx.x];
    const Inst1D in = insts[tk.inst];
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N;
    if (N <= WAVE_EIG_MAX_N) return;               // the wave eigensolver's matrix
    const int64_t NN = (int64_t)N * N;
    cd* base = ws + lay[layout_layer(in.L, tk.layer)];
    cd *S0 = base, *S1 = base + NN, *S2 = base + 2 * NN;
    cd* lam = base + 4 * NN;
    EigScratch s{base + 4 * NN + N, base + 4 * NN + 2 * N};
    const int info = wg_eig<QC>(N, S0, lam, S1, S2, s, max_iter, kEps);
    if (threadIdx.x == 0) task_info[blockIdx.x] = info;
}

// One wave per layer matrix (orders up to WAVE_EIG_MAX_N): four layers per workgroup.
template <int QC>
__global__ __launch_bounds__(NT) void k_eig1d_wave(int n_tasks, const Inst1D* __restrict__ insts,
                                                   const LayerTask* __restrict__ tasks,
                                                   const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                   int* __restrict__ task_info, int max_iter) {
    __shared__ double s_rot[NWAVE][3 * (64 * QC + 1)];
    const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * NWAVE + wave_id());   // wave-uniform
    if (t >= n_tasks || task_info[t] != 0) return;
    const LayerTask tk = tasks[t];
    const Inst1D in = insts[tk.inst];
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N;
    if (N > WAVE_EIG_MAX_N) return;                // the workgroup eigensolver's matrix
    const int64_t NN = (int64_t)N * N;
    cd* base = ws + lay[layout_layer(in.L, tk.layer)];
    cd *S0 = base, *S1 = base + NN, *S2 = base + 2 * NN;
    cd* lam = base + 4 * NN;
    EigScratch s{base + 4 * NN + N, base + 4 * NN + 2 * N};
    const int info = wave_eig<QC>(N, S0, lam, S1, S2, s, s_rot[wave_id()], max_iter, kEps);
    if (lane_id() == 0) task_info[t] = info;
}

// QcLaunch::what: 0 = k_eig1d (workgroup per matrix), 1 = k_eig1d_wave (wave per matrix)
SRCWA_QC_DISPATCH(planar_qc, QcLaunch)
#ifdef SRCWA_QC_ONLY
void SRCWA_QC_FN(planar_qc)(const QcLaunch& a) {
    constexpr int QC = SRCWA_QC_ONLY;
    const Inst1D* insts = static_cast<const Inst1D*>(a.insts);
    if (a.what == 0)
        k_eig1d<QC><<<a.grid, NT, 0, a.stream>>>(insts, a.tasks, a.layouts, a.ws, a.task_info, a.max_iter);
    else if constexpr (QC <= 4)
        k_eig1d_wave<QC><<<a.grid, NT, 0, a.stream>>>(a.n_items, insts, a.tasks, a.layouts, a.ws, a.task_info,
                                                      a.max_iter);
}
#endif

// ------------------------------------------------------------------------------------------------
#if SRCWA_MAIN_TU
__global__ __launch_bounds__(NT) void k_modes1d(const Inst1D* __restrict__ insts, const LayerTask* __restrict__ tasks,
                                                const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                const int* __restrict__ task_info, const LayerOps* __restrict__ descs,
                                                cd* __restrict__ ops) {
    if (task_info[blockIdx.x] != 0) return;
    const LayerTask tk = tasks[blockIdx.x];
    const Inst1D in = insts[tk.inst];
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N;
    const int64_t NN = (int64_t)N * N;
    cd* base = ws + lay[layout_layer(in.L, tk.layer)];
    cd *W = base, *V = base + NN;
    // TM: [[1/eps]], kept by the assembly in S3 or held by the task's shared operator set
    const cd* S3 = shares_ops(tk, in) && in.pol != 0 ? planar_ops_at(ops + descs[tk.op].off, N, in.pol).Ti : base + 3 * NN;
    cd* lam = base + 4 * NN;
    cd* q = ws + lay[layout_q(tk.layer + 1)];
    cd* basis = base + 4 * NN + N;
    for (int j = threadIdx.x; j < N; j += NT) {
        const cd qq = branch_q(csqrt(lam[j]));
        q[j] = qq;
        basis[j] = is_zero(qq) ? cd{1.0, 0.0} : qq;
    }
    __syncthreads();
    if (in.pol == 0) {
        wg_scale_cols(N, N, W, N, basis, V, N);
    } else {
        wg_gemm<GEMM_SET, false>(N, N, N, S3, N, W, N, V, N);
        wg_scale_cols(N, N, V, N, basis, V, N);
    }
}
#endif

// ------------------------------------------------------------------------------------------------
// FIELDS: the plan keeps the near-field capture (PLAN_FIELDS).  The two forms are instantiated in units of their
// own (part 1 without, part 2 with the capture), so plans without fields run the kernel with no capture code.
#if SRCWA_IN_PART(1) || SRCWA_IN_PART(2)
template <bool FIELDS>
__global__ SMAT_KERNEL_BOUNDS_OF(SRCWA_SMAT_OCC_PLANAR) void k_smatrix1d(const Inst1D* __restrict__ insts, const double* __restrict__ stacks,
                                                  const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                  const int* __restrict__ task_info, Out1D out) {
    const Inst1D in = insts[blockIdx.x];
    const double* st = stacks + in.stack_off;
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, M = in.M, L = in.L;
    const double k0 = 2.0 * kPi / in.wavelength;
    const cd nI = cd{st[1], st[2]}, nII = cd{st[3], st[4]};
    cd* R = reinterpret_cast<cd*>(out.R) + in.out_off;
    cd* T = reinterpret_cast<cd*>(out.T) + in.out_off;
    double* DEr = out.DEr + in.out_off;
    double* DEt = out.DEt + in.out_off;

    // failures of assembly or eigensolve (first nonzero eig info in layer order)
    int eig_info = 0, failed = 0;
    double max_res = 0.0, max_berr = 0.0;
    int n_solve = 0, n_eig = 0;
    const bool diag = in.flags & PLAN_DIAGNOSTICS;
    for (int t = 0; t < in.n_tasks; ++t) {
        const int v = task_info[in.first_task + t];
        if (v < 0) failed = 1;
        else if (v > 0 && eig_info == 0) eig_info = v;
        const TaskDiag d = out.task_diag[in.first_task + t];
        n_solve += d.n_solve;
        n_eig += d.n_eig;
        if (diag) {
            max_berr = max_keep_nan(max_berr, d.berr);
            max_res = max_keep_nan(max_res, d.eig_res);
        }
    }
    int status = (failed || eig_info) ? ST_NUMERICAL : ST_OK;
    SmatCtx c;
    c.N = N;
    c.fail = 0;
    c.diag = diag;
    c.mixed = in.flags & PLAN_MIXED;
    if (status == ST_OK) {
        __shared__ double s_thick[SMAT_MAX_LAYERS];
        LayerCursor cur(st);
        bool pure = true;
        c.bump = Bump(ws, in.ws_lo, in.ws_hi);
        // regions: 0 = incidence, j+1 = layer j, L+1 = exit
        auto region = [&](int r) -> Region {
            Region g;
            g.q = ws + lay[layout_q(r)];
            g.p = ws + lay[layout_p(r)];
            g.w = nullptr;
            g.W = g.V = g.X = nullptr;
            if (r == 0 || r == L + 1) {
                g.analytic = true;
                g.fzb = false;
                return g;
            }
            const int64_t b = lay[layout_layer(L, r - 1)];
            if (b < 0) {
                g.analytic = true;
                g.fzb = true;
            } else {
                g.analytic = false;
                g.fzb = true;
                g.W = ws + b;
                g.V = ws + b + (int64_t)N * N;
            }
            return g;
        };
        for (int j = 0; j < L; ++j) {
            const LayerView lv = cur.next();
            if (threadIdx.x == 0) s_thick[j] = lv.d;
            pure = pure && lv.homogeneous;
        }
        __syncthreads();
        const int cols[1] = {M};
        cd* fields = FIELDS ? reinterpret_cast<cd*>(out.fields) + in.field_off : nullptr;
        run_recursion(c, L, k0, s_thick, region, 1, cols, R, T, fields, pure, M, N);
        status = c.fail ? (c.fail == 2 ? ST_WORKSPACE : ST_NUMERICAL) : ST_OK;
        if (FIELDS && status == ST_OK) {
            // the capture is usable when every mode set and amplitude is finite (rcwa1d.hpp:496-508)
            int bad = 0;
            for (int64_t e = threadIdx.x; e < 2 * (int64_t)L * N; e += NT) bad |= !finite(fields[e]);
            for (int r = 0; r < L + 2; ++r) {
                const Region g = region(r);
                for (int i = threadIdx.x; i < N; i += NT) bad |= !finite(g.q[i]) || !finite(g.p[i]);
                if (!g.analytic)
                    for (int64_t e = threadIdx.x; e < (int64_t)N * N; e += NT) bad |= !finite(g.W[e]) || !finite(g.V[e]);
            }
            bad = block_max_int(bad);
            if (threadIdx.x == 0) out.field_ok[blockIdx.x] = !bad;
        }
        if (status == ST_OK) {
            __syncthreads();
            if (pure)
                for (int j = threadIdx.x; j < N; j += NT) {
                    if (j == M) continue;
                    if (!finite(R[j])) R[j] = cd{0.0, 0.0};
                    if (!finite(T[j])) T[j] = cd{0.0, 0.0};
                }
            __syncthreads();
            const cd* qI = ws + lay[layout_q(0)];
            const cd* qII = ws + lay[layout_q(L + 1)];
            const cd den{-k0 * qI[M].y, k0 * qI[M].x};
            int bad = 0;
            for (int j = threadIdx.x; j < N; j += NT) {
                const cd r = R[j], t = T[j];
                if (!finite(r) || !finite(t)) bad = 1;
                const cd kIz{-k0 * qI[j].y, k0 * qI[j].x};
                const cd kIIz{-k0 * qII[j].y, k0 * qII[j].x};
                const double der = norm2(r) * cdiv(kIz, den).x;
                cd factor;
                if (in.pol == 0) factor = cdiv(kIIz, den);
                else factor = cdiv(cdiv(kIIz, nII * nII), cdiv(den, nI * nI));
                const double det = norm2(t) * factor.x;
                if (!isfinite(der) || !isfinite(det)) bad = 1;
                DEr[j] = der;
                DEt[j] = det;
            }
            if (block_max_int(bad)) status = ST_NUMERICAL;
        }
    }
    __syncthreads();
    if (status != ST_OK) {
        const double nan = __builtin_nan("");
        for (int j = threadIdx.x; j < N; j += NT) {
            R[j] = cd{nan, nan};
            T[j] = cd{nan, nan};
            DEr[j] = nan;
            DEt[j] = nan;
        }
    }
    if (threadIdx.x == 0) {
        if (FIELDS && status != ST_OK) out.field_ok[blockIdx.x] = 0;
        out.status[blockIdx.x] = status;
        out.eig_info[blockIdx.x] = eig_info;
        double* dg = out.diag + (int64_t)DIAG_WORDS * blockIdx.x;
        dg[0] = max_res;
        dg[1] = max_keep_nan(max_berr, c.max_berr);
        dg[2] = n_eig;
        dg[3] = n_solve + c.n_solve;
        dg[4] = 0.0;
        c.refine_record(dg);
    }
}
#endif
#if SRCWA_IN_PART(1)
void smatrix1d_launch(int grid, hipStream_t s, const Inst1D* insts, const double* stacks, const int64_t* layouts, cd* ws, const int* task_info, Out1D out) {
    k_smatrix1d<false><<<grid, NT, 0, s>>>(insts, stacks, layouts, ws, task_info, out);
}
#endif
#if SRCWA_IN_PART(2)
void smatrix1d_fields_launch(int grid, hipStream_t s, const Inst1D* insts, const double* stacks, const int64_t* layouts, cd* ws, const int* task_info, Out1D out) {
    k_smatrix1d<true><<<grid, NT, 0, s>>>(insts, stacks, layouts, ws, task_info, out);
}
#endif

// ------------------------------------------------------------------------------------------------
// the S-matrix kernel is compiled in units of its own
void smatrix1d_launch(int grid, hipStream_t s, const Inst1D* insts, const double* stacks, const int64_t* layouts, cd* ws, const int* task_info, Out1D out);
void smatrix1d_fields_launch(int grid, hipStream_t s, const Inst1D* insts, const double* stacks, const int64_t* layouts, cd* ws, const int* task_info, Out1D out);

#if SRCWA_MAIN_TU
// ------------------------------------------------------------------------------------------------
// Near fields of one planar solve (PlanarFieldSampler::sample, rcwa1d_fields.hpp:107-189): one workgroup per
// depth group.  From the capture the S-matrix kernel kept (a+, a- per layer, R, T) and the modes still in
// the workspace (q, p per region; W, V per grating layer): the harmonic amplitudes S, U of every depth,
// then the six components (Ex, Ey, Ez, Hx, Hy, Hz; H is Z0 H), out[c][i + iz N].  TM needs Ez = [[eps]]^-1
// Kx S, with [[eps]] of a grating layer rebuilt and factored here by the assembly's own routines.
__global__ __launch_bounds__(NT) void k_fields1d(const Inst1D* __restrict__ inst, const double* __restrict__ stacks,
                                                 const int64_t* __restrict__ layouts, const cd* __restrict__ ws,
                                                 const cd* __restrict__ fields, const cd* __restrict__ R,
                                                 const cd* __restrict__ T, const FieldGroup1D* __restrict__ groups,
                                                 const int* __restrict__ cols, const double* __restrict__ zs, int nz,
                                                 cd* __restrict__ out, cd* __restrict__ scratch) {
    const FieldGroup1D g = groups[blockIdx.x];
    const Inst1D in = *inst;
    const double* st = stacks + in.stack_off;
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, M = in.M, L = in.L, cnt = g.count;
    const int64_t NC = (int64_t)N * cnt, NZ = (int64_t)N * nz;
    const double k0 = 2.0 * kPi / in.wavelength;
    const cd* kx = ws + lay[LAYOUT_KX];
    cd* Sm = scratch + g.scratch;
    cd* Um = Sm + NC;
    cd* Ez = Um + NC;
    const cd *Sf = Sm, *Uf = Um;
    cd eps_scale{0.0, 0.0};          // Ez = eps_scale Kx S outside grating layers
    if (g.region >= 0) {
        const int j = g.region, r = j + 1;
        const cd* q = ws + lay[layout_q(r)];
        const cd* p = ws + lay[layout_p(r)];
        const int64_t lb = lay[layout_layer(L, j)];
        const cd* ap = fields + in.field_off + (int64_t)j * N;
        const cd* am = fields + in.field_off + (int64_t)(L + j) * N;
        const double d = g.d;
        for (int64_t e = threadIdx.x; e < NC; e += NT) {
            const int i = (int)(e % N), col = (int)(e / N);
            const double z = fmax(0.0, fmin(d, zs[cols[g.first + col]] - g.z_top));
            const cd a = ap[i], b = am[i], qi = q[i];
            cd S, U;
            if (qi.x == 0.0 && qi.y == 0.0) {
                // the finite-zero basis: linear in z, distances formed first (rcwa1d_fields.hpp:124-131)
                const double den = 1.0 + 0.5 * k0 * d;
                const double from_bottom = (1.0 + k0 * (d - z)) / den;
                const double from_top = (1.0 + k0 * z) / den;
                S = from_bottom * a + from_top * b;
                U = cdiv(-a, den) + cdiv(b, den);
            } else {
                // each term decays from its own anchor; a zero amplitude takes no exponential
                const bool az = a.x == 0.0 && a.y == 0.0, bz = b.x == 0.0 && b.y == 0.0;
                const cd pp = az ? cd{0.0, 0.0} : a * cexp((-k0 * z) * qi);
                const cd mm = bz ? cd{0.0, 0.0} : b * cexp((-k0 * (d - z)) * qi);
                S = pp + mm;
                U = mm - pp;
            }
            Sm[e] = S;
            Um[e] = lb < 0 ? U * p[i] : U;
        }
        __syncthreads();
        if (lb >= 0) {
            const cd* W = ws + lb;
            const cd* V = W + (int64_t)N * N;
            cd* S2 = Ez + NC;
            cd* U2 = S2 + NC;
            wg_gemm<GEMM_SET, false>(N, cnt, N, W, N, Sm, N, S2, N);
            wg_gemm<GEMM_SET, false>(N, cnt, N, V, N, Um, N, U2, N);
            Sf = S2;
            Uf = U2;
        } else {
            LayerCursor cur(st);
            LayerView lv = cur.next();
            for (int k = 0; k < j; ++k) lv = cur.next();
            eps_scale = cdiv(cd{1.0, 0.0}, lv.n * lv.n);
        }
    } else {
        const int r = g.region == -1 ? 0 : L + 1;
        const cd* p = ws + lay[layout_p(r)];
        const cd* qI = ws + lay[layout_q(0)];
        const cd* qII = ws + lay[layout_q(L + 1)];
        const bool top = g.region == -1;
        // kz = i k0 q of the two outer media (the field metadata's kzI, kzII)
        const cd kz0{-k0 * qI[M].y, k0 * qI[M].x};
        for (int64_t e = threadIdx.x; e < NC; e += NT) {
            const int i = (int)(e % N), col = (int)(e / N);
            const double z = zs[cols[g.first + col]];
            const cd zero{0.0, 0.0};
            cd pp = zero, rr = zero, tt = zero;
            if (top) {
                const cd kzi{-k0 * qI[i].y, k0 * qI[i].x};
                if (i == M) pp = cexp(cd{-kz0.y, kz0.x} * z);
                if (R[i].x != 0.0 || R[i].y != 0.0) rr = R[i] * cexp(cd{kzi.y, -kzi.x} * z);
            } else if (T[i].x != 0.0 || T[i].y != 0.0) {
                const cd kzt{-k0 * qII[i].y, k0 * qII[i].x};
                tt = T[i] * cexp(cd{-kzt.y, kzt.x} * (z - g.z_bottom));
            }
            Sm[e] = pp + rr + tt;
            Um[e] = p[i] * ((rr - pp) - tt);
        }
        __syncthreads();
        const cd n = top ? cd{st[1], st[2]} : cd{st[3], st[4]};
        eps_scale = cdiv(cd{1.0, 0.0}, n * n);
    }
    // a polarization's three other components are zero (make_harmonic_fields starts every component at zero)
    const cd zero{0.0, 0.0};
    if (in.pol == 0) {
        for (int64_t e = threadIdx.x; e < NC; e += NT) {
            const int i = (int)(e % N), iz = cols[g.first + (int)(e / N)];
            const cd S = Sf[e], U = Uf[e];
            out[0 * NZ + i + (int64_t)iz * N] = zero;
            out[1 * NZ + i + (int64_t)iz * N] = S;
            out[2 * NZ + i + (int64_t)iz * N] = zero;
            out[3 * NZ + i + (int64_t)iz * N] = cd{-U.y, U.x};
            out[4 * NZ + i + (int64_t)iz * N] = zero;
            out[5 * NZ + i + (int64_t)iz * N] = cd{-kx[i].x * S.x, -kx[i].x * S.y};
        }
        return;
    }
    for (int64_t e = threadIdx.x; e < NC; e += NT) {
        const int i = (int)(e % N);
        const cd v = kx[i].x * Sf[e];
        Ez[e] = g.tm_grating ? v : v * eps_scale;
    }
    __syncthreads();
    if (g.tm_grating) {
        // [[eps]] of the layer, LU-factored as in the TM assembly (rcwa1d.hpp:455-462)
        cd* E = Ez + 3 * NC;
        int* piv = reinterpret_cast<int*>(E + (int64_t)N * N);
        cd* coef = E + (int64_t)N * N + ((N + 15) / 16) * 16;
        cd* F = coef + ((4 * M + 1 + 15) / 16) * 16;
        LayerCursor cur(st);
        LayerView lv = cur.next();
        for (int k = 0; k < g.region; ++k) lv = cur.next();
        layer_coefficients(lv, st[0], M, false, coef, F);
        toeplitz(coef, N, E);
        wg_lu(N, E, N, piv);
        wg_lu_solve(N, E, N, piv, Ez, N, cnt);
    }
    for (int64_t e = threadIdx.x; e < NC; e += NT) {
        const int i = (int)(e % N), iz = cols[g.first + (int)(e / N)];
        const cd U = Uf[e];
        out[0 * NZ + i + (int64_t)iz * N] = cd{U.y, -U.x};
        out[1 * NZ + i + (int64_t)iz * N] = zero;
        out[2 * NZ + i + (int64_t)iz * N] = Ez[e];
        out[3 * NZ + i + (int64_t)iz * N] = zero;
        out[4 * NZ + i + (int64_t)iz * N] = Sf[e];
        out[5 * NZ + i + (int64_t)iz * N] = zero;
    }
}

// A harmonic field set on an x grid (reconstruct_field_grid, rcwa_fields.hpp:326-355): component c at
// (x_ix, z_iz) = sum_j exp(-i (kxm_j x_ix + ky y)) H[c][j + iz N], x_ix = ix period / nx; out[c][ix + iz nx].
// kxm_j = k0 kx_j from the first N words of the instance's layout block (planar kx, conical geometry);
// kyy = ky y (0 for planar solves).
__global__ __launch_bounds__(NT) void k_fieldmap1d(const Inst1D* __restrict__ inst, const double* __restrict__ stacks,
                                                   const int64_t* __restrict__ layouts, const cd* __restrict__ ws,
                                                   const cd* __restrict__ H, int nz, int nx, double kyy,
                                                   cd* __restrict__ out) {
    const Inst1D in = *inst;
    const int N = in.N;
    const int64_t cells = (int64_t)nx * nz;
    const int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (e >= 6 * cells) return;
    const int c = (int)(e / cells);
    const int64_t k = e % cells;
    const int ix = (int)(k % nx), iz = (int)(k / nx);
    const double period = stacks[in.stack_off];
    const double k0 = 2.0 * kPi / in.wavelength;
    const cd* kx = ws + layouts[in.layout_off + LAYOUT_KX];
    const double x = (static_cast<double>(ix) / nx) * period;
    const cd* h = H + (int64_t)c * N * nz + (int64_t)iz * N;
    cd s{0.0, 0.0};
    for (int j = 0; j < N; ++j) {
        double sn, cs;
        sincos((k0 * kx[j].x) * x + kyy, &sn, &cs);
        s = s + cd{cs, -sn} * h[j];
    }
    out[e] = s;
}

void launch_fields1d(hipStream_t s, int n_groups, const Inst1D* inst, const double* stacks, const int64_t* layouts,
                     const cd* ws, const cd* fields, const cd* R, const cd* T, const FieldGroup1D* groups,
                     const int* cols, const double* zs, int nz, cd* out, cd* scratch) {
    k_fields1d<<<n_groups, NT, 0, s>>>(inst, stacks, layouts, ws, fields, R, T, groups, cols, zs, nz, out, scratch);
}
void launch_fieldmap1d(hipStream_t s, const Inst1D* inst, const double* stacks, const int64_t* layouts, const cd* ws,
                       const cd* H, int nz, int nx, double kyy, cd* out) {
    const int64_t total = 6 * (int64_t)nx * nz;
    k_fieldmap1d<<<(unsigned)((total + NT - 1) / NT), NT, 0, s>>>(inst, stacks, layouts, ws, H, nz, nx, kyy, out);
}

// host launchers
void launch_planar(hipStream_t s, int n_inst, int n_tasks, const Inst1D* insts, const LayerTask* tasks,
                   const double* stacks, const int64_t* layouts, cd* ws, int* task_info, Out1D out, int max_iter, int min_n, int max_n,
                   int part, int flags, hipEvent_t* ev, const OpsArgs& ops) {
    const bool diag = flags & PLAN_DIAGNOSTICS;
    // part 1: shared operators (first chunk), setup, assembly, eigensolves (events ev[0..2]); part 2:
    // modes, S-matrix, efficiencies (events ev[3], ev[4]); the host may run the parts of different chunks
    // on two streams
    if (part & 1) {
        if (ev) hipEventRecord(ev[0], s);
        if (ops.compute && ops.n > 0)
            k_ops1d<<<ops.n, NT, 0, s>>>(ops.descs, static_cast<const Inst1D*>(ops.insts), stacks, ops.ops, ops.info);
        k_setup1d<<<n_inst, NT, 0, s>>>(insts, stacks, layouts, ws);
        if (n_tasks > 0)
            k_assemble1d<<<n_tasks, NT, 0, s>>>(insts, tasks, stacks, layouts, ws, task_info, out.task_diag, ops.descs,
                                                ops.ops, ops.info);
        if (ev) hipEventRecord(ev[1], s);
        // each matrix goes to the wave or the workgroup eigensolver by its own order, so a result does not
        // depend on what else is in the batch
        QcLaunch a{};
        a.stream = s;
        a.n_items = n_tasks;
        a.insts = insts;
        a.tasks = tasks;
        a.layouts = layouts;
        a.ws = ws;
        a.task_info = task_info;
        a.max_iter = max_iter;
        if (n_tasks > 0 && min_n <= WAVE_EIG_MAX_N) {
            a.what = 1;
            a.grid = (n_tasks + NWAVE - 1) / NWAVE;
            planar_qc(wave_columns_per_lane(WAVE_EIG_MAX_N), a);
        }
        if (n_tasks > 0 && max_n > WAVE_EIG_MAX_N) {
            a.what = 0;
            a.grid = n_tasks;
            planar_qc(eig_columns_per_thread(max_n), a);
        }
        if (n_tasks > 0 && diag) k_eigres1d<<<n_tasks, NT, 0, s>>>(insts, tasks, layouts, ws, task_info, out.task_diag);
        if (ev) hipEventRecord(ev[2], s);
    }
    if (part & 2) {
        if (ev) hipEventRecord(ev[3], s);
        if (n_tasks > 0) k_modes1d<<<n_tasks, NT, 0, s>>>(insts, tasks, layouts, ws, task_info, ops.descs, ops.ops);
        if (out.fields) smatrix1d_fields_launch(n_inst, s, insts, stacks, layouts, ws, task_info, out);
        else smatrix1d_launch(n_inst, s, insts, stacks, layouts, ws, task_info, out);
        if (ev) hipEventRecord(ev[4], s);
    }
}
#endif
