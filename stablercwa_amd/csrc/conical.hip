// Conical-incidence RCWA on gfx950 (the reference's solve_conical, rcwa1d_conical.hpp:678-982).
//
//   k_setup_conical     one workgroup per instance: kx_m, ky, the per-order azimuths (cos, sin),
//                       kz and flux factors of the outer media, and the analytic (s, t)-basis modes
//                       of homogeneous regions                   (rcwa1d_conical.hpp:195-283, 724-747)
//   k_assemble_conical  one workgroup per grating layer: [[eps]]^-1, C = [[1/eps]]^-1, B C and
//                       A = Kx^2 - [[eps]]                       (rcwa1d_conical.hpp:302-336)
//   k_eig_conical       one workgroup per N x N eigensolve: A and B C of every grating layer
//   k_modes_conical     one workgroup per grating layer: the degeneracy test, then the 2N modes
//                       W, V in the (s, t) basis                 (rcwa1d_conical.hpp:341-411, 826-848)
//   k_smatrix_conical   one workgroup per instance: the recursion at order 2N with the s and p
//                       incident columns, Jones matrices and channel efficiencies
//                                                                (rcwa1d_conical.hpp:852-956)
// Zero-thickness layers do not take part (the host lists only layers with d > 0).  A layer whose
// reduced eigenvalues come within 1e-2 |ky|^1.5 of zero takes the reference's Schur-basis
// propagation (rcwa1d_conical.hpp:413-502): Schur form of the 2N-order P Q, triangular square
// root, V = Q Z G^-1, and a dense exp(-k0 d G) propagation block.
#include "batch.hpp"
#include "eig.hpp"
#include "wave_eig.hpp"
#include "rcwa_common.hpp"

// Conical layout table (int64 cd offsets), see conical_layout_* in batch.hpp.

// kz = sqrt(n^2 - kx^2 - ky^2) with Im >= 0 (Re >= 0 when Im ~ 0); exact sign rules for real media
// (rcwa1d_conical.hpp:193-203).
DEV cd longitudinal(cd n, double kx, double ky) {
    const cd z = n * n - cd{kx * kx + 0.0, 0.0} - cd{ky * ky, 0.0};
    if (z.y == 0.0) return z.x >= 0.0 ? cd{sqrt(z.x), 0.0} : cd{0.0, sqrt(-z.x)};
    cd kz = csqrt(z);
    if (kz.y < 0.0) kz = -kz;
    if (fabs(kz.y) < 1e-12 && kz.x < 0.0) kz = -kz;
    return kz;
}
// zero order: (n - nI)(n + nI) + (nI cos theta)^2 (rcwa1d_conical.hpp:264-272)
DEV cd longitudinal0(cd n, double nI, double cosine) {
    const double t = nI * cosine;
    const cd z = (n - cd{nI, 0.0}) * (n + cd{nI, 0.0}) + cd{t * t, 0.0};
    if (z.y == 0.0) return z.x >= 0.0 ? cd{sqrt(z.x), 0.0} : cd{0.0, sqrt(-z.x)};
    cd kz = csqrt(z);
    if (kz.y < 0.0) kz = -kz;
    if (fabs(kz.y) < 1e-12 && kz.x < 0.0) kz = -kz;
    return kz;
}

// s branch (w, v) = (1, q), t branch (w, v) = (kz / n, -i n), q = -i kz (rcwa1d_conical.hpp:273-291)
DEV void analytic_region(cd n, const cd* kx, double ky, int N, int M, double nI, double cosine, cd* q, cd* w, cd* v) {
    for (int j = threadIdx.x; j < N; j += NT) {
        const cd kz = j == M ? longitudinal0(n, nI, cosine) : longitudinal(n, kx[j].x, ky);
        const cd qq{kz.y, -kz.x};
        q[j] = qq;
        q[N + j] = qq;
        w[j] = cd{1.0, 0.0};
        v[j] = qq;
        w[N + j] = cdiv(kz, n);
        v[N + j] = cd{n.y, -n.x};
    }
}

struct ConicalGeo {
    cd *kx, *cs, *sn, *kzI, *kzII, *flux;   // cs, sn, flux (r in .x, t in .y) real
};
DEV ConicalGeo geo_of(cd* ws, const int64_t* lay, int N) {
    cd* g = ws + lay[0];
    return ConicalGeo{g, g + N, g + 2 * N, g + 3 * N, g + 4 * N, g + 5 * N};
}

// a grating layer's slots (order N): A, BC, IE, C, Z1, X1, Z2, X2 (N x N); W, V (2N x 2N);
// lambda1, lambda2, ort1, tmp1, ort2, tmp2, pivots (N)
struct ConicalSlots {
    cd *A, *BC, *IE, *C, *Z1, *X1, *Z2, *X2, *W, *V, *lam1, *lam2, *ort1, *tmp1, *ort2, *tmp2;
    int* piv;
};
DEV ConicalSlots slots_of(cd* base, int N) {
    const int64_t NN = (int64_t)N * N;
    ConicalSlots s;
    s.A = base;
    s.BC = base + NN;
    s.IE = base + 2 * NN;
    s.C = base + 3 * NN;
    s.Z1 = base + 4 * NN;
    s.X1 = base + 5 * NN;
    s.Z2 = base + 6 * NN;
    s.X2 = base + 7 * NN;
    s.W = base + 8 * NN;
    s.V = base + 12 * NN;
    cd* v = base + 16 * NN;
    s.lam1 = v;
    s.lam2 = v + N;
    s.ort1 = v + 2 * N;
    s.tmp1 = v + 3 * N;
    s.ort2 = v + 4 * N;
    s.tmp2 = v + 5 * N;
    s.piv = reinterpret_cast<int*>(v + 6 * N);
    return s;
}

DEV LayerView layer_at(const double* st, int index) {
    LayerCursor cur(st);
    LayerView lv = cur.next();
    for (int j = 0; j < index; ++j) lv = cur.next();
    return lv;
}

// Field plans (PLAN_FIELDS): the Schur generator G (2N x 2N) of a degenerate layer, kept for sampling at any
// depth, after the layer's slots, diagnostics copies and FFT samples (the host sizes the slot to match).
DEV cd* schur_generator_keep(cd* ws, const int64_t* lay, const Inst1D& in, int layer, const LayerView& lv) {
    const int64_t NN = (int64_t)in.N * in.N;
    const int64_t fft = (!lv.homogeneous && lv.method == 1) ? 2 * (int64_t)fft_samples_for(in.M, lv.samples) : 0;
    return ws + lay[conical_slot(in.L, layer)] + conical_slot_words(in.N) + ((in.flags & PLAN_DIAGNOSTICS) ? 4 * NN : 0) + fft;
}

// A conical operator set (layout: conical_ops_words): [[eps]], [[eps]]^-1, C = [[1/eps]]^-1.
struct ConicalOps {
    cd *E, *IE, *C, *Z, *coef, *F;
    int* piv;
};
DEV ConicalOps conical_ops_at(cd* base, int N) {
    const int64_t m = ops_align16((int64_t)N * N);
    ConicalOps o;
    o.E = base;
    o.IE = base + m;
    o.C = base + 2 * m;
    o.Z = base + 3 * m;
    o.piv = reinterpret_cast<int*>(base + 4 * m);
    o.coef = base + 4 * m + ops_align16(N);
    o.F = o.coef + ops_align16(2 * (int64_t)N);
    return o;
}
DEV bool shares_ops(const LayerTask& tk, const Inst1D& in) { return tk.op >= 0 && !(in.flags & PLAN_DIAGNOSTICS); }
// a task's slots, with [[eps]]^-1 and C read from its shared operator set when it has one
DEV ConicalSlots task_slots(const LayerTask& tk, const Inst1D& in, cd* ws, const int64_t* lay,
                            const LayerOps* descs, cd* ops) {
    ConicalSlots s = slots_of(ws + lay[conical_slot(in.L, tk.layer)], in.N);
    if (shares_ops(tk, in)) {
        const ConicalOps o = conical_ops_at(ops + descs[tk.op].off, in.N);
        s.IE = o.IE;
        s.C = o.C;
    }
    return s;
}

// ------------------------------------------------------------------------------------------------
#if SRCWA_MAIN_TU
__global__ __launch_bounds__(NT) void k_setup_conical(const Inst1D* __restrict__ insts, const double* __restrict__ stacks,
                                                      const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                      double* __restrict__ geometry) {
    const Inst1D in = insts[blockIdx.x];
    const double* st = stacks + in.stack_off;
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, M = in.M, L = in.L;
    const double period = st[0];
    const cd nI = cd{st[1], st[2]}, nII = cd{st[3], st[4]};
    const double theta = in.theta_deg * kPi / 180.0, phi = in.phi_deg * kPi / 180.0;
    const double nIr = nI.x;
    const double sine = nIr * sin(theta);
    const double ky = sine * sin(phi);
    const double cth = cos(theta);
    const double flux_unit = nIr * cth;
    ConicalGeo g = geo_of(ws, lay, N);
    for (int j = threadIdx.x; j < N; j += NT) {
        const double kxj = sine * cos(phi) - (j - M) * in.wavelength / period;
        const double kt = hypot(kxj, ky);
        g.kx[j] = cd{kxj, 0.0};
        if (kt <= 1e-14) {
            g.cs[j] = cd{cos(phi), 0.0};
            g.sn[j] = cd{sin(phi), 0.0};
        } else {
            g.cs[j] = cd{kxj / kt, 0.0};
            g.sn[j] = cd{ky / kt, 0.0};
        }
        const cd kzI = j == M ? longitudinal0(nI, nIr, cth) : longitudinal(nI, kxj, ky);
        const cd kzII = j == M ? longitudinal0(nII, nIr, cth) : longitudinal(nII, kxj, ky);
        g.kzI[j] = kzI;
        g.kzII[j] = kzII;
        g.flux[j] = cd{kzI.x / flux_unit, kzII.x / flux_unit};
        if (geometry) {
            // the diffraction-plane azimuth of each order (rcwa1d_conical.hpp:735-742)
            double* o = geometry + GEO_WORDS * (in.out_off + j);
            o[0] = kxj;
            o[1] = kzI.x;
            o[2] = kzI.y;
            o[3] = kzII.x;
            o[4] = kzII.y;
            o[5] = kzI.x / flux_unit;
            o[6] = kzII.x / flux_unit;
            o[7] = kt <= 1e-14 ? in.phi_deg : atan2(ky, kxj) * 180.0 / kPi;
        }
    }
    if (threadIdx.x == 0) g.kx[6 * N] = cd{ky, 0.0};
    __syncthreads();
    analytic_region(nI, g.kx, ky, N, M, nIr, cth, ws + lay[conical_q(0)], ws + lay[conical_w(0)], ws + lay[conical_v(0)]);
    analytic_region(nII, g.kx, ky, N, M, nIr, cth, ws + lay[conical_q(L + 1)], ws + lay[conical_w(L + 1)],
                    ws + lay[conical_v(L + 1)]);
    for (int j = 0; j < L; ++j) {
        if (lay[conical_slot(L, j)] >= 0) continue;
        const LayerView lv = layer_at(st, (int)lay[conical_orig(L, j)]);
        const cd n = lv.homogeneous ? lv.n : cd{lv.vals[0], lv.vals[1]};
        analytic_region(n, g.kx, ky, N, M, nIr, cth, ws + lay[conical_q(j + 1)], ws + lay[conical_w(j + 1)],
                        ws + lay[conical_v(j + 1)]);
    }
    __syncthreads();
}
#endif

// ------------------------------------------------------------------------------------------------
#if SRCWA_MAIN_TU
// The wavelength-invariant operators of one grating layer, with the arithmetic of the per-task form in
// k_assemble_conical (rcwa1d_conical.hpp:302-336): [[eps]], its inverse and the inverse of [[1/eps]].
__global__ __launch_bounds__(NT) void k_ops_conical(const LayerOps* __restrict__ descs, const Inst1D* __restrict__ insts,
                                                    const double* __restrict__ stacks, cd* __restrict__ ops,
                                                    int* __restrict__ ops_info) {
    const LayerOps d = descs[blockIdx.x];
    const Inst1D in = insts[d.inst];
    const double* st = stacks + in.stack_off;
    const int N = in.N, M = in.M;
    const LayerView lv = layer_at(st, d.layer);
    const ConicalOps o = conical_ops_at(ops + d.off, N);
    int fail = layer_coefficients(lv, st[0], M, false, o.coef, o.F);
    toeplitz(o.coef, N, o.E);                  // [[eps]]
    wg_copy(N, N, o.E, N, o.Z, N);
    fail |= wg_lu(N, o.Z, N, o.piv);
    wg_identity(N, o.IE, N);
    if (!fail) wg_lu_solve(N, o.Z, N, o.piv, o.IE, N, N);            // [[eps]]^-1
    fail |= layer_coefficients(lv, st[0], M, true, o.coef, o.F);
    toeplitz(o.coef, N, o.Z);                  // [[1/eps]]
    fail |= wg_lu(N, o.Z, N, o.piv);
    wg_identity(N, o.C, N);
    if (!fail) wg_lu_solve(N, o.Z, N, o.piv, o.C, N, N);             // C = [[1/eps]]^-1
    if (threadIdx.x == 0) ops_info[blockIdx.x] = fail;
}

__global__ __launch_bounds__(NT) void k_assemble_conical(const Inst1D* __restrict__ insts,
                                                         const LayerTask* __restrict__ tasks,
                                                         const double* __restrict__ stacks,
                                                         const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                         int* __restrict__ task_info, TaskDiag* __restrict__ task_diag,
                                                         const LayerOps* __restrict__ descs, cd* __restrict__ ops,
                                                         const int* __restrict__ ops_info) {
    const LayerTask tk = tasks[blockIdx.x];
    const Inst1D in = insts[tk.inst];
    const double* st = stacks + in.stack_off;
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, M = in.M, L = in.L;
    const int64_t NN = (int64_t)N * N;
    const LayerView lv = layer_at(st, (int)lay[conical_orig(L, tk.layer)]);
    const ConicalGeo g = geo_of(ws, lay, N);
    // with a shared operator set, [[eps]]^-1 and C are read from it and [[eps]] comes from its copy
    const bool shared = shares_ops(tk, in);
    const ConicalSlots s = task_slots(tk, in, ws, lay, descs, ops);
    const cd* E = shared ? conical_ops_at(ops + descs[tk.op].off, N).E : s.A;
    cd* coef = s.lam1;                         // lambda1 .. tmp2 hold the 4M+1 coefficients
    // with diagnostics: D0, D1 keep A and B C for the eigen residuals, D2, D3 are solve scratch
    const bool diag = in.flags & PLAN_DIAGNOSTICS;
    cd* D0 = ws + lay[conical_slot(L, tk.layer)] + conical_slot_words(N);
    cd *D1 = D0 + NN, *D2 = D0 + 2 * NN, *D3 = D0 + 3 * NN;
    cd* F = D0 + (diag ? 4 * NN : 0);          // FFT layers: 2 nfft sample scratch
    double berr = 0.0;
    int fail = shared ? ops_info[tk.op] : 0;
    if (!shared) {
        fail |= layer_coefficients(lv, st[0], M, false, coef, F);
        toeplitz(coef, N, s.A);                // [[eps]]
        wg_copy(N, N, s.A, N, s.Z1, N);
        fail |= wg_lu(N, s.Z1, N, s.piv);
        wg_identity(N, s.IE, N);
        if (!fail) wg_lu_solve(N, s.Z1, N, s.piv, s.IE, N, N);      // [[eps]]^-1
        if (diag && !fail) {
            wg_identity(N, D2, N);
            berr = max_keep_nan(berr, wg_backward_error(N, s.A, s.IE, D2, N, D3));
        }
        fail |= layer_coefficients(lv, st[0], M, true, coef, F);
        toeplitz(coef, N, s.Z2);               // [[1/eps]]
        if (diag) wg_copy(N, N, s.Z2, N, D0, N);
        fail |= wg_lu(N, s.Z2, N, s.piv);
        wg_identity(N, s.C, N);
        if (!fail) wg_lu_solve(N, s.Z2, N, s.piv, s.C, N, N);       // C = [[1/eps]]^-1
        if (diag && !fail) berr = max_keep_nan(berr, wg_backward_error(N, D0, s.C, D2, N, D3));
    }
    // B = Kx [[eps]]^-1 Kx - I, then B C
    for (int64_t e = threadIdx.x; e < NN; e += NT) {
        const int i = (int)(e % N), j = (int)(e / N);
        cd v = (g.kx[i].x * s.IE[e]) * g.kx[j].x;
        if (i == j) v = v - cd{1.0, 0.0};
        s.X1[e] = v;
    }
    __syncthreads();
    wg_gemm<GEMM_SET, false>(N, N, N, s.X1, N, s.C, N, s.BC, N);
    // A = Kx^2 - [[eps]]
    for (int64_t e = threadIdx.x; e < NN; e += NT) {
        const int i = (int)(e % N), j = (int)(e / N);
        cd v = -E[e];
        if (i == j) v = v + cd{g.kx[i].x * g.kx[i].x, 0.0};
        s.A[e] = v;
    }
    __syncthreads();
    int bad = 0;
    for (int64_t e = threadIdx.x; e < NN; e += NT) bad |= !finite(s.A[e]) || !finite(s.BC[e]);
    bad = block_max_int(bad);
    if (diag) {
        wg_copy(N, N, s.A, N, D0, N);
        wg_copy(N, N, s.BC, N, D1, N);
    }
    if (threadIdx.x == 0) {
        task_info[4 * blockIdx.x] = (fail || bad) ? -1 : 0;
        // two dense solves (the inverses of [[eps]] and [[1/eps]], rcwa1d_conical.hpp:302-336), two
        // eigensolves (A and B C); a Schur-basis layer adds one more, counted from its modes status
        task_diag[blockIdx.x] = TaskDiag{0.0, berr, 2, 2};
    }
}

// task_diag[t].eig_res = the larger residual of the task's two eigensolves
__global__ __launch_bounds__(NT) void k_fold_res(int n_tasks, const double* __restrict__ item_res,
                                                 TaskDiag* __restrict__ task_diag) {
    const int t = blockIdx.x * NT + threadIdx.x;
    if (t < n_tasks) task_diag[t].eig_res = max_keep_nan(item_res[2 * t], item_res[2 * t + 1]);
}

// Eigen-pair residuals of a layer's two eigensolves (A and B C) from the matrices kept in D0 and D1
// (diagnostics only); blockIdx.x = 2 task + which, Z1 / Z2 are scratch (free until the modes kernel).
__global__ __launch_bounds__(NT) void k_eigres_conical(const Inst1D* __restrict__ insts,
                                                       const LayerTask* __restrict__ tasks,
                                                       const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                       const int* __restrict__ task_info,
                                                       double* __restrict__ item_res) {
    const int t = blockIdx.x >> 1, which = blockIdx.x & 1;
    const LayerTask tk = tasks[t];
    const Inst1D in = insts[tk.inst];
    if (!(in.flags & PLAN_DIAGNOSTICS)) return;
    const int N = in.N;
    const int64_t NN = (int64_t)N * N;
    const int64_t* lay = layouts + in.layout_off;
    cd* base = ws + lay[conical_slot(in.L, tk.layer)];
    const ConicalSlots s = slots_of(base, N);
    cd* D = base + conical_slot_words(N) + which * NN;
    double res = __builtin_nan("");
    if (task_info[4 * t] == 0 && task_info[4 * t + 1 + which] == 0)
        res = wg_eig_residual(N, D, which ? s.BC : s.A, which ? s.lam2 : s.lam1, which ? s.Z2 : s.Z1);
    if (threadIdx.x == 0) item_res[blockIdx.x] = res;
}
#endif


// ------------------------------------------------------------------------------------------------
template <int QC>
__global__ EIG_KERNEL_BOUNDS_OF(SRCWA_EIG_OCC_CONICAL) void k_eig_conical(const Inst1D* __restrict__ insts, const LayerTask* __restrict__ tasks,
                                                    const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                    int* __restrict__ task_info, int max_iter, const int* __restrict__ order) {
    // the two eigensolves of a layer task are neighbouring blocks; tasks largest first
    const int b = blockIdx.x >> 1, which = blockIdx.x & 1;
    const int t = order ? order[b] : b;
    if (task_info[4 * t] != 0) return;
    const LayerTask tk = tasks[t];
    const Inst1D in = insts[tk.inst];
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N;
    if (N <= WAVE_EIG_MAX_N) return;               // the wave eigensolver's matrix
    const ConicalSlots s = slots_of(ws + lay[conical_slot(in.L, tk.layer)], N);
    int info;
    if (which == 0) info = wg_eig<QC>(N, s.A, s.lam1, s.Z1, s.X1, EigScratch{s.ort1, s.tmp1}, max_iter, kEps);
    else info = wg_eig<QC>(N, s.BC, s.lam2, s.Z2, s.X2, EigScratch{s.ort2, s.tmp2}, max_iter, kEps);
    if (threadIdx.x == 0) task_info[4 * t + 1 + which] = info;
}


// One wave per eigenproblem (two per layer: A and B C), four per workgroup.
template <int QC>
__global__ __launch_bounds__(NT) void k_eig_conical_wave(int n_items, const Inst1D* __restrict__ insts,
                                                         const LayerTask* __restrict__ tasks,
                                                         const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                         int* __restrict__ task_info, int max_iter) {
    __shared__ double s_rot[NWAVE][3 * (64 * QC + 1)];
    const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * NWAVE + wave_id());   // wave-uniform
    if (item >= n_items) return;
    const int t = item >> 1, which = item & 1;
    if (task_info[4 * t] != 0) return;
    const LayerTask tk = tasks[t];
    const Inst1D in = insts[tk.inst];
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N;
    if (N > WAVE_EIG_MAX_N) return;                // the workgroup eigensolver's matrix
    const ConicalSlots s = slots_of(ws + lay[conical_slot(in.L, tk.layer)], N);
    int info;
    if (which == 0)
        info = wave_eig<QC>(N, s.A, s.lam1, s.Z1, s.X1, EigScratch{s.ort1, s.tmp1}, s_rot[wave_id()], max_iter, kEps);
    else
        info = wave_eig<QC>(N, s.BC, s.lam2, s.Z2, s.X2, EigScratch{s.ort2, s.tmp2}, s_rot[wave_id()], max_iter, kEps);
    if (lane_id() == 0) task_info[4 * t + 1 + which] = info;
}

// ------------------------------------------------------------------------------------------------
// Schur-basis modes of a degenerate layer (rcwa1d_conical.hpp:413-502), order n = 2N.
// Scratch E1..E4 (n x n each) come from the instance's S-matrix workspace, which is idle until the
// S-matrix kernel runs; each grating layer of an instance owns its own quarter-sized share.

// max_j sum_i |a(i, j)|
DEV double wg_norm1(int n, const cd* a) {
    double best = 0.0;
    for (int j = wave_id(); j < n; j += NWAVE) {
        double sum = 0.0;
        for (int i = lane_id(); i < n; i += 64) sum += cabs(a[i + (int64_t)j * n]);
        sum = wave_sum(sum);
        best = fmax(best, sum);
    }
    return block_max(best);
}

// G = sqrt(T) for upper triangular T, in place; fails on a zero or non-finite root or denominator.
DEV int wg_triangular_root(int n, cd* g) {
    int bad = 0;
    for (int i = threadIdx.x; i < n; i += NT) {
        const cd r = branch_q(csqrt(g[i + (int64_t)i * n]));
        if (is_zero(r) || !finite(r)) bad = 1;
        g[i + (int64_t)i * n] = r;
    }
    if (block_max_int(bad)) return 1;
    for (int gap = 1; gap < n; ++gap) {
        for (int i = threadIdx.x; i + gap < n; i += NT) {
            const int j = i + gap;
            cd v = g[i + (int64_t)j * n];
            for (int k = i + 1; k < j; ++k) v = v - g[i + (int64_t)k * n] * g[k + (int64_t)j * n];
            const cd den = g[i + (int64_t)i * n] + g[j + (int64_t)j * n];
            if (is_zero(den)) bad = 1;
            g[i + (int64_t)j * n] = cdiv(v, den);
        }
        __syncthreads();
    }
    if (block_max_int(bad)) return 1;
    int nf = 0;
    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) nf |= !finite(g[e]);
    return block_max_int(nf);
}

// exp(a) of an upper triangular a by scaled Taylor series and squaring, with the diagonal replaced
// by exact exponentials after every stage (rcwa1d_conical.hpp:441-462).  a, e, t1, t2: n x n;
// diag_a[i] = a(i, i).  Returns 1 on failure; e holds the result.
DEV int wg_triangular_exp(int n, cd* a, const cd* diag_a, cd* e, cd* t1, cd* t2) {
    const int64_t nn = (int64_t)n * n;
    const double size = wg_norm1(n, a);
    if (!isfinite(size)) return 1;
    const int scaling = size > 0.25 ? max(0, (int)ceil(log2(size) + 2.0)) : 0;
    if (scaling > 1020) return 1;
    const double sc = ldexp(1.0, -scaling);
    for (int64_t x = threadIdx.x; x < nn; x += NT) {
        a[x] = a[x] * sc;
        const int i = (int)(x % n), j = (int)(x / n);
        e[x] = cd{i == j ? 1.0 : 0.0, 0.0};
        t1[x] = e[x];
    }
    __syncthreads();
    bool converged = n == 0;
    for (int k = 1; k <= 40 && !converged; ++k) {
        wg_gemm<GEMM_SET, false>(n, n, n, t1, n, a, n, t2, n);
        const double inv = 1.0 / k;
        int bad = 0;
        for (int64_t x = threadIdx.x; x < nn; x += NT) {
            const cd term = t2[x] * cd{inv, 0.0};
            t1[x] = term;
            e[x] = e[x] + term;
            bad |= !finite(e[x]) || !finite(term);
        }
        if (block_max_int(bad)) return 1;
        const double nt = wg_norm1(n, t1), ne = wg_norm1(n, e);
        converged = nt <= kEps * fmax(1.0, ne);
    }
    if (!converged) return 1;
    for (int i = threadIdx.x; i < n; i += NT) e[i + (int64_t)i * n] = cexp(diag_a[i] * sc);
    __syncthreads();
    for (int k = 0; k < scaling; ++k) {
        wg_gemm<GEMM_SET, false>(n, n, n, e, n, e, n, t2, n);
        const double f = ldexp(1.0, k + 1 - scaling);
        int bad = 0;
        for (int64_t x = threadIdx.x; x < nn; x += NT) {
            const int i = (int)(x % n), j = (int)(x / n);
            e[x] = i == j ? cexp(diag_a[i] * f) : t2[x];
            bad |= !finite(e[x]);
        }
        if (block_max_int(bad)) return 1;
    }
    return 0;
}

// The degenerate layer's Schur-basis modes: returns 1 on numerical failure, else writes q (2N),
// W (= Z), V (= Q Z G^-1) and the propagation X = exp(-k0 d G) (into the Z1..X2 slots).
template <int QC>
DEV int schur_modes(const Inst1D& in, const double* st, const int64_t* lay, cd* ws, int layer, int ordinal,
                    const ConicalGeo& g, const ConicalSlots& s, cd* q, double ky, int max_iter) {
    const int N = in.N, M = in.M, n = 2 * N;
    const int64_t NN = (int64_t)N * N, nn = (int64_t)n * n;
    cd* E1 = ws + in.ws_lo + (int64_t)ordinal * 4 * nn;
    cd* E2 = E1 + nn;
    cd* E3 = E2 + nn;
    cd* E4 = E3 + nn;
    // A = Kx^2 - [[eps]] again (the eigensolve overwrote it): into the Z1 slot
    const LayerView lv = layer_at(st, (int)lay[conical_orig(in.L, layer)]);
    cd* F = ws + lay[conical_slot(in.L, layer)] + conical_slot_words(N) + ((in.flags & PLAN_DIAGNOSTICS) ? 4 * NN : 0);
    if (layer_coefficients(lv, st[0], M, false, s.lam1, F)) return 1;
    toeplitz(s.lam1, N, s.Z1);
    for (int64_t e = threadIdx.x; e < NN; e += NT) {
        const int i = (int)(e % N), j = (int)(e / N);
        cd v = -s.Z1[e];
        if (i == j) v = v + cd{g.kx[i].x * g.kx[i].x, 0.0};
        s.Z1[e] = v;
    }
    __syncthreads();
    // P and Q (rcwa1d_conical.hpp:470-481) into E1, E2
    for (int64_t e = threadIdx.x; e < NN; e += NT) {
        const int i = (int)(e % N), j = (int)(e / N);
        const double kxi = g.kx[i].x, kxj = g.kx[j].x;
        const cd id{i == j ? 1.0 : 0.0, 0.0};
        const cd ie = s.IE[e];
        E1[i + (int64_t)j * n] = (ky * kxi) * ie;
        E1[i + (int64_t)(N + j) * n] = id - (kxi * ie) * kxj;
        E1[N + i + (int64_t)j * n] = (ky * ky) * ie - id;
        E1[N + i + (int64_t)(N + j) * n] = ((-ky) * ie) * kxj;
        E2[i + (int64_t)j * n] = cd{ky * kxi, 0.0} * id;
        E2[i + (int64_t)(N + j) * n] = -s.Z1[e];
        E2[N + i + (int64_t)j * n] = cd{ky * ky, 0.0} * id - s.C[e];
        E2[N + i + (int64_t)(N + j) * n] = cd{-ky * kxi, 0.0} * id;
    }
    __syncthreads();
    // P <- re P rv, Q <- rv Q re with the per-order 2 x 2 rotations
    //   re = [[-sn, cs], [cs, sn]], rv = [[-cs, -sn], [-sn, cs]] (blocks over orders i, N+i)
    for (int64_t e = threadIdx.x; e < nn; e += NT) {              // E3 = re P, E4 = rv Q
        const int r = (int)(e % n), c = (int)(e / n);
        const int i = r % N;
        const double cs = g.cs[i].x, sn = g.sn[i].x;
        const cd p0 = E1[i + (int64_t)c * n], p1 = E1[N + i + (int64_t)c * n];
        const cd q0 = E2[i + (int64_t)c * n], q1 = E2[N + i + (int64_t)c * n];
        if (r < N) {
            E3[e] = (-sn) * p0 + cs * p1;
            E4[e] = (-cs) * q0 + (-sn) * q1;
        } else {
            E3[e] = cs * p0 + sn * p1;
            E4[e] = (-sn) * q0 + cs * q1;
        }
    }
    __syncthreads();
    for (int64_t e = threadIdx.x; e < nn; e += NT) {              // E1 = (re P) rv, E2 = (rv Q) re
        const int r = (int)(e % n), c = (int)(e / n);
        const int j = c % N;
        const double cs = g.cs[j].x, sn = g.sn[j].x;
        const cd x0 = E3[r + (int64_t)j * n], x1 = E3[r + (int64_t)(N + j) * n];
        const cd y0 = E4[r + (int64_t)j * n], y1 = E4[r + (int64_t)(N + j) * n];
        if (c < N) {
            E1[e] = x0 * (-cs) + x1 * (-sn);
            E2[e] = y0 * (-sn) + y1 * cs;
        } else {
            E1[e] = x0 * (-sn) + x1 * cs;
            E2[e] = y0 * cs + y1 * sn;
        }
    }
    __syncthreads();
    // H = P Q, Schur form H = Z T Z^H (max_iter 100, rcwa1d_conical.hpp:487-489): T in E3, Z in W
    wg_gemm<GEMM_SET, false>(n, n, n, E1, n, E2, n, E3, n);
    int bad = 0;
    for (int64_t e = threadIdx.x; e < nn; e += NT) bad |= !finite(E3[e]);
    if (block_max_int(bad)) return 1;
    cd* wv = s.lam1;              // eigenvalues and the two scratch vectors: lambda1 .. tmp2 (6N = 3n entries)
    if (wg_eig<QC>(n, E3, wv, s.W, E4, EigScratch{wv + n, wv + 2 * n}, max_iter, kEps, false) != 0) return 1;
    // G = sqrt(T) with a residual check ||G G - T||_1 <= 1e-10 ||T||_1
    wg_copy(n, n, E3, n, E1, n);
    if (wg_triangular_root(n, E3)) return 1;
    wg_gemm<GEMM_SET, false>(n, n, n, E3, n, E3, n, E4, n);
    for (int64_t e = threadIdx.x; e < nn; e += NT) E4[e] = E4[e] - E1[e];
    __syncthreads();
    const double scale = wg_norm1(n, E1), err = wg_norm1(n, E4);
    if (!isfinite(scale) || !isfinite(err) || err > 1e-10 * fmax(scale, 2.2250738585072014e-308)) return 1;
    for (int i = threadIdx.x; i < n; i += NT) q[i] = E3[i + (int64_t)i * n];
    if (in.flags & PLAN_FIELDS) {
        cd* keep = schur_generator_keep(ws, lay, in, layer, lv);
        for (int64_t e = threadIdx.x; e < nn; e += NT) keep[e] = E3[e];
    }
    __syncthreads();
    // V = Q Z G^-1 (right triangular solve, one thread per row)
    wg_gemm<GEMM_SET, false>(n, n, n, E2, n, s.W, n, s.V, n);
    for (int i = threadIdx.x; i < n; i += NT) {
        for (int j = 0; j < n; ++j) {
            cd v = s.V[i + (int64_t)j * n];
            for (int k = 0; k < j; ++k) v = v - s.V[i + (int64_t)k * n] * E3[k + (int64_t)j * n];
            s.V[i + (int64_t)j * n] = cdiv(v, E3[j + (int64_t)j * n]);
        }
    }
    __syncthreads();
    bad = 0;
    for (int64_t e = threadIdx.x; e < nn; e += NT) bad |= !finite(s.V[e]);
    if (block_max_int(bad)) return 1;
    // X = exp(-k0 d G) into the Z1..X2 slots (4 N^2 = n^2, contiguous)
    const double k0 = 2.0 * kPi / in.wavelength;
    const double dk = -k0 * lv.d;
    for (int64_t e = threadIdx.x; e < nn; e += NT) E1[e] = E3[e] * cd{dk, 0.0};
    for (int i = threadIdx.x; i < n; i += NT) E2[i] = E3[i + (int64_t)i * n] * cd{dk, 0.0};
    __syncthreads();
    cd* diag_a = s.lam1;          // n entries: lambda1 .. tmp1 are free now
    for (int i = threadIdx.x; i < n; i += NT) diag_a[i] = E2[i];
    __syncthreads();
    return wg_triangular_exp(n, E1, diag_a, s.Z1, E2, E4);
}

// ------------------------------------------------------------------------------------------------
template <int QC>
__global__ __launch_bounds__(NT) void k_modes_conical(const Inst1D* __restrict__ insts, const LayerTask* __restrict__ tasks,
                                                      const double* __restrict__ stacks,
                                                      const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                      int* __restrict__ task_info, const LayerOps* __restrict__ descs,
                                                      cd* __restrict__ ops) {
    const int t = blockIdx.x;
    if (task_info[4 * t] != 0 || task_info[4 * t + 1] != 0 || task_info[4 * t + 2] != 0) return;
    const LayerTask tk = tasks[t];
    const Inst1D in = insts[tk.inst];
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, n = 2 * N;
    const int64_t NN = (int64_t)N * N;
    const ConicalGeo g = geo_of(ws, lay, N);
    const ConicalSlots s = task_slots(tk, in, ws, lay, descs, ops);
    cd* q = ws + lay[conical_q(tk.layer + 1)];
    // ---- degeneracy test (rcwa1d_conical.hpp:826-848)
    double lm = INFINITY;
    for (int j = threadIdx.x; j < N; j += NT) lm = fmin(lm, fmin(cabs(s.lam1[j]), cabs(s.lam2[j])));
    const double local_min = -block_max(-lm);
    const double kyy = ws[lay[0] + 6 * N].x;
    const double ky_scale = fabs(kyy) * sqrt(fabs(kyy));
    if (kyy != 0.0 && local_min < 1e-2 * ky_scale) {
        const int fail = schur_modes<QC>(in, stacks + in.stack_off, lay, ws, tk.layer, t - in.first_task, g, s, q, kyy, 100);
        if (threadIdx.x == 0) task_info[4 * t + 3] = fail ? -1 : 3;
        return;
    }
    // ---- general region (rcwa1d_conical.hpp:341-411)
    const cd ky2{kyy * kyy, 0.0};
    for (int j = threadIdx.x; j < N; j += NT) {
        q[j] = branch_q(csqrt(s.lam1[j] + ky2));
        q[N + j] = branch_q(csqrt(s.lam2[j] + ky2));
    }
    // cw = C W2 -> X1; coupling = [[eps]]^-1 (ky Kx C W2) -> X2 (W2 lives in BC, W1 in A)
    wg_gemm<GEMM_SET, false>(N, N, N, s.C, N, s.BC, N, s.X1, N);
    if (kyy != 0.0) {
        for (int64_t e = threadIdx.x; e < NN; e += NT) {
            const int m = (int)(e % N);
            s.Z1[e] = (kyy * g.kx[m].x) * s.X1[e];
        }
        __syncthreads();
        wg_gemm<GEMM_SET, false>(N, N, N, s.IE, N, s.Z1, N, s.X2, N);
    }
    const int64_t nn = (int64_t)n;
    for (int64_t e = threadIdx.x; e < (int64_t)N * N; e += NT) {
        const int m = (int)(e % N), j = (int)(e / N);
        const double csm = g.cs[m].x, snm = g.sn[m].x;
        {   // first group: Ex = 0, Ey = W1 q1^2, hx = -W1 l1 q1, hy = -ky Kx W1 q1
            const cd q1 = q[j];
            const cd fy = q1 * q1, gx = -(s.lam1[j] * q1), gy = (-kyy) * q1;
            const cd w1 = s.A[m + (int64_t)j * N];
            const cd ey = w1 * fy, hx = w1 * gx, hy = (g.kx[m].x * w1) * gy;
            s.W[m + j * nn] = csm * ey;
            s.W[N + m + j * nn] = snm * ey;
            s.V[m + j * nn] = -(csm * hx) - snm * hy;
            s.V[N + m + j * nn] = csm * hy - snm * hx;
        }
        {   // second group: Ex = W2 l2, Ey = ky [[eps]]^-1 Kx C W2, hx = 0, hy = -C W2 q2
            const cd q2 = q[N + j], l2 = s.lam2[j];
            const cd ex = s.BC[m + (int64_t)j * N] * l2;
            const cd ey = kyy != 0.0 ? s.X2[m + (int64_t)j * N] : cd{0.0, 0.0};
            const cd hy = -(s.X1[m + (int64_t)j * N] * q2);
            const int64_t col = (int64_t)(N + j) * nn;
            s.W[m + col] = csm * ey - snm * ex;
            s.W[N + m + col] = csm * ex + snm * ey;
            s.V[m + col] = -(snm * hy);
            s.V[N + m + col] = csm * hy;
        }
    }
    __syncthreads();
    // normalise every column of [W; V] to unit 2-norm; a zero or non-finite column fails the layer
    int bad = 0;
    for (int col = wave_id(); col < n; col += NWAVE) {
        cd* wc = s.W + col * nn;
        cd* vc = s.V + col * nn;
        double sum = 0.0;
        for (int i = lane_id(); i < n; i += 64) sum += norm2(wc[i]) + norm2(vc[i]);
        sum = wave_sum(sum);
        const double scale = sqrt(sum);
        if (!(scale > 0.0) || !isfinite(scale)) {
            bad = 1;
            continue;
        }
        for (int i = lane_id(); i < n; i += 64) {
            wc[i] = cdiv(wc[i], scale);
            vc[i] = cdiv(vc[i], scale);
        }
    }
    bad = block_max_int(bad);
    if (threadIdx.x == 0) task_info[4 * t + 3] = bad ? -1 : 0;
}

// QcLaunch::what: 0 = k_eig_conical, 1 = k_eig_conical_wave, 2 = k_modes_conical
SRCWA_QC_DISPATCH(conical_qc, QcLaunch)
#ifdef SRCWA_QC_ONLY
void SRCWA_QC_FN(conical_qc)(const QcLaunch& a) {
    constexpr int QC = SRCWA_QC_ONLY;
    const Inst1D* insts = static_cast<const Inst1D*>(a.insts);
    if (a.what == 0) {
        k_eig_conical<QC><<<a.grid, NT, 0, a.stream>>>(insts, a.tasks, a.layouts, a.ws, a.task_info, a.max_iter, a.order);
    } else if (a.what == 1) {
        if constexpr (QC <= 4)
            k_eig_conical_wave<QC><<<a.grid, NT, 0, a.stream>>>(a.n_items, insts, a.tasks, a.layouts, a.ws,
                                                                a.task_info, a.max_iter);
    } else {
        k_modes_conical<QC><<<a.grid, NT, 0, a.stream>>>(insts, a.tasks, a.stacks, a.layouts, a.ws, a.task_info,
                                                         a.op_descs, a.ops);
    }
}
#endif

// ------------------------------------------------------------------------------------------------
// FIELDS: the plan keeps the near-field capture (PLAN_FIELDS): per layer a+ and a- (2N x 2, columns s and p
// incidence), then the outer responses up and down (2N x 2 each), at Inst1D::field_off.  The two forms are
// instantiated in units of their own (part 1 without, part 2 with the capture).
#if SRCWA_IN_PART(1) || SRCWA_IN_PART(2)
template <bool FIELDS>
__global__ SMAT_KERNEL_BOUNDS_OF(SRCWA_SMAT_OCC_CONICAL) void k_smatrix_conical(const Inst1D* __restrict__ insts,
                                                        const double* __restrict__ stacks,
                                                        const int64_t* __restrict__ layouts, cd* __restrict__ ws,
                                                        const int* __restrict__ task_info, Out1D out) {
    const Inst1D in = insts[blockIdx.x];
    const double* st = stacks + in.stack_off;
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, M = in.M, L = in.L, n = 2 * N;
    const double k0 = 2.0 * kPi / in.wavelength;
    const ConicalGeo geo = geo_of(ws, lay, N);
    cd* Jr = reinterpret_cast<cd*>(out.R) + 4 * in.out_off;
    cd* Jt = reinterpret_cast<cd*>(out.T) + 4 * in.out_off;
    double* DEr = out.DEr + 4 * in.out_off;
    double* DEt = out.DEt + 4 * in.out_off;

    int eig_info = 0, failed = 0;
    double max_res = 0.0, max_berr = 0.0;
    int n_solve = 0, n_eig = 0, n_schur = 0;
    const bool diag = in.flags & PLAN_DIAGNOSTICS;
    SmatCtx c;
    c.N = n;
    c.fail = 0;
    c.diag = diag;
    c.mixed = in.flags & PLAN_MIXED;
    for (int t = 0; t < in.n_tasks; ++t) {
        const int* ti = task_info + 4 * (in.first_task + t);
        if (ti[0] < 0 || ti[3] < 0) failed = 1;
        for (int w = 1; w <= 2; ++w)
            if (ti[w] != 0 && eig_info == 0) eig_info = ti[w];
        const TaskDiag d = out.task_diag[in.first_task + t];
        n_solve += d.n_solve;
        n_eig += d.n_eig + (ti[3] == 3 ? 1 : 0);
        n_schur += ti[3] == 3 ? 1 : 0;
        if (diag) {
            max_berr = max_keep_nan(max_berr, d.berr);
            max_res = max_keep_nan(max_res, d.eig_res);
        }
    }
    int status = (failed || eig_info) ? ST_NUMERICAL : ST_OK;
    // a uniform layer with an exactly grazing order fails (pure film stacks: unexcited orders excepted)
    bool pure = true;
    for (int j = 0; j < L; ++j) pure = pure && lay[conical_slot(L, j)] < 0;
    if (status == ST_OK) {
        int zero = 0;
        for (int j = 0; j < L; ++j) {
            if (lay[conical_slot(L, j)] >= 0) continue;
            const cd* q = ws + lay[conical_q(j + 1)];
            for (int k = threadIdx.x; k < N; k += NT)
                if (is_zero(q[k]) && (!pure || k == M)) zero = 1;
        }
        if (block_max_int(zero)) status = ST_NUMERICAL;
    }
    __shared__ double s_thick[SMAT_MAX_LAYERS];
    if (status == ST_OK) {
        for (int j = 0; j < L; ++j) {
            const LayerView lv = layer_at(st, (int)lay[conical_orig(L, j)]);
            if (threadIdx.x == 0) s_thick[j] = lv.d;
        }
        __syncthreads();
        c.bump = Bump(ws, in.ws_lo, in.ws_hi);
        auto region = [&](int r) -> Region {
            Region g;
            g.fzb = false;
            g.X = nullptr;
            g.q = ws + lay[conical_q(r)];
            if (r == 0 || r == L + 1 || lay[conical_slot(L, r - 1)] < 0) {
                g.analytic = true;
                g.w = ws + lay[conical_w(r)];
                g.p = ws + lay[conical_v(r)];
                g.W = g.V = nullptr;
            } else {
                const ConicalSlots sl = slots_of(ws + lay[conical_slot(L, r - 1)], N);
                g.analytic = false;
                g.w = nullptr;
                g.p = nullptr;
                g.W = sl.W;
                g.V = sl.V;
                // a Schur-basis layer propagates with its dense exp(-k0 d G) (task status 3)
                int ord = 0;
                for (int jj = 0; jj < r - 1; ++jj) ord += lay[conical_slot(L, jj)] >= 0;
                if (task_info[4 * (in.first_task + ord) + 3] == 3) g.X = sl.Z1;
            }
            return g;
        };
        cd* upm = c.bump.keep(2 * (int64_t)n);
        cd* downm = c.bump.keep(2 * (int64_t)n);
        c.check();
        const int cols[2] = {M, N + M};
        cd* fields = FIELDS ? reinterpret_cast<cd*>(out.fields) + in.field_off : nullptr;
        if (!c.fail) run_recursion(c, L, k0, s_thick, region, 2, cols, upm, downm, fields, pure, M, N);
        status = c.fail ? (c.fail == 2 ? ST_WORKSPACE : ST_NUMERICAL) : ST_OK;
        if (FIELDS && status == ST_OK) {
            // the outer responses (before the reflected p sign) and the capture's validity
            // (rcwa1d_conical.hpp:232-241, 957-970); a Schur-basis layer's generator G is kept by its modes kernel
            cd* resp = fields + 4 * (int64_t)L * n;
            int bad = 0;
            for (int e = threadIdx.x; e < 2 * n; e += NT) {
                const bool keep = !pure || (e % n) % N == M;
                resp[e] = keep ? upm[e] : cd{0.0, 0.0};
                resp[2 * n + e] = keep ? downm[e] : cd{0.0, 0.0};
                bad |= !finite(resp[e]) || !finite(resp[2 * n + e]);
            }
            for (int64_t e = threadIdx.x; e < 4 * (int64_t)L * n; e += NT) bad |= !finite(fields[e]);
            cd* marks = fields + 4 * (int64_t)(L + 1) * n;       // 1 for a Schur-basis layer, then its G is kept
            for (int r = 0; r < L + 2; ++r) {
                const Region g = region(r);
                if (r >= 1 && r <= L) {
                    if (threadIdx.x == 0) marks[r - 1] = cd{g.X != nullptr ? 1.0 : 0.0, 0.0};
                    if (g.X != nullptr) {
                        const LayerView lv = layer_at(st, (int)lay[conical_orig(L, r - 1)]);
                        const cd* G = schur_generator_keep(ws, lay, in, r - 1, lv);
                        for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) bad |= !finite(G[e]);
                    }
                }
                for (int i = threadIdx.x; i < n; i += NT) {
                    bad |= !finite(g.q[i]);
                    if (g.analytic) bad |= !finite(g.w[i]) || !finite(g.p[i]);
                }
                if (!g.analytic)
                    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) bad |= !finite(g.W[e]) || !finite(g.V[e]);
            }
            bad = block_max_int(bad);
            if (threadIdx.x == 0) out.field_ok[blockIdx.x] = !bad;
        }
        if (status == ST_OK) {
            int bad = 0;
            for (int j = threadIdx.x; j < N; j += NT) {
                // outer media: the s amplitude is Es; the t amplitude is Ep downwards and -Ep upwards
                cd jr[4] = {upm[j], upm[n + j], -upm[N + j], -upm[n + N + j]};
                cd jt[4] = {downm[j], downm[n + j], downm[N + j], downm[n + N + j]};
                if (pure && j != M)
                    for (int k = 0; k < 4; ++k) {
                        if (!finite(jr[k])) jr[k] = cd{0.0, 0.0};
                        if (!finite(jt[k])) jt[k] = cd{0.0, 0.0};
                    }
                const double fr = geo.flux[j].x, ft = geo.flux[j].y;
                for (int k = 0; k < 4; ++k) {
                    if (!finite(jr[k]) || !finite(jt[k])) bad = 1;
                    Jr[4 * j + k] = jr[k];
                    Jt[4 * j + k] = jt[k];
                    const double er = fr * norm2(jr[k]), et = ft * norm2(jt[k]);
                    if (!isfinite(er) || !isfinite(et)) bad = 1;
                    DEr[4 * j + k] = er;
                    DEt[4 * j + k] = et;
                }
            }
            if (block_max_int(bad)) status = ST_NUMERICAL;
        }
    }
    __syncthreads();
    if (status != ST_OK) {
        const double nan = __builtin_nan("");
        for (int j = threadIdx.x; j < 4 * N; j += NT) {
            Jr[j] = cd{nan, nan};
            Jt[j] = cd{nan, nan};
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
        dg[4] = n_schur;
        c.refine_record(dg);
    }
}
#endif
#if SRCWA_IN_PART(1)
void smatrix_conical_launch(int grid, hipStream_t s, const Inst1D* insts, const double* stacks, const int64_t* layouts, cd* ws, const int* task_info, Out1D out) {
    k_smatrix_conical<false><<<grid, NT, 0, s>>>(insts, stacks, layouts, ws, task_info, out);
}
#endif
#if SRCWA_IN_PART(2)
void smatrix_conical_fields_launch(int grid, hipStream_t s, const Inst1D* insts, const double* stacks, const int64_t* layouts, cd* ws, const int* task_info, Out1D out) {
    k_smatrix_conical<true><<<grid, NT, 0, s>>>(insts, stacks, layouts, ws, task_info, out);
}
#endif

// ------------------------------------------------------------------------------------------------
// the S-matrix kernel is compiled in units of its own
void smatrix_conical_launch(int grid, hipStream_t s, const Inst1D* insts, const double* stacks, const int64_t* layouts, cd* ws, const int* task_info, Out1D out);
void smatrix_conical_fields_launch(int grid, hipStream_t s, const Inst1D* insts, const double* stacks, const int64_t* layouts, cd* ws, const int* task_info, Out1D out);

#if SRCWA_MAIN_TU
// ------------------------------------------------------------------------------------------------
// Near fields of one conical solve for one incident Jones vector (ConicalFieldSampler::sample,
// rcwa1d_conical_fields.hpp:158-236): one workgroup per depth group.  The captured 2N x 2 amplitudes are
// combined with the (normalized) Jones vector (j0, j1), propagated from their anchors, mapped to the
// tangential fields F = W (p + m), G = V (m - p) (analytic regions: w, v scaling), rotated from the (s, t)
// basis of each order to x, y, and Ez = [[eps]]^-1 (i (ky hx - kx hy)) (a uniform region: / eps).
// out[c][i + iz N], c = Ex, Ey, Ez, Hx, Hy, Hz.
__global__ __launch_bounds__(NT) void k_fields_conical(const Inst1D* __restrict__ inst, const double* __restrict__ stacks,
                                                       const int64_t* __restrict__ layouts, const cd* __restrict__ ws,
                                                       const cd* __restrict__ fields, const FieldGroup1D* __restrict__ groups,
                                                       const int* __restrict__ cols, const double* __restrict__ zs, int nz,
                                                       cd j0, cd j1, cd* __restrict__ out, cd* __restrict__ scratch) {
    const FieldGroup1D g = groups[blockIdx.x];
    const Inst1D in = *inst;
    const double* st = stacks + in.stack_off;
    const int64_t* lay = layouts + in.layout_off;
    const int N = in.N, M = in.M, L = in.L, n = 2 * N, cnt = g.count;
    const int64_t Nc = (int64_t)n * cnt, NC = (int64_t)N * cnt, NZ = (int64_t)N * nz;
    const double k0 = 2.0 * kPi / in.wavelength;
    const ConicalGeo geo = geo_of(const_cast<cd*>(ws), lay, N);
    const double theta = in.theta_deg * kPi / 180.0, phi = in.phi_deg * kPi / 180.0;
    const double ky = st[1] * sin(theta) * sin(phi);
    cd* P = scratch + g.scratch;
    cd* Q = P + Nc;
    cd* F = Q + Nc;
    cd* G = F + Nc;
    cd* E1 = G + Nc;
    cd* E2 = E1 + NC;
    const cd* Ff = P;
    const cd* Gf = Q;
    const cd* IE = nullptr;           // [[eps]]^-1 of a grating region
    cd eps_scale{0.0, 0.0};
    if (g.region >= 0) {
        const int j = g.region, r = j + 1;
        const cd* q = ws + lay[conical_q(r)];
        const int64_t sb = lay[conical_slot(L, j)];
        const cd* ap = fields + in.field_off + (int64_t)j * 2 * n;
        const cd* am = fields + in.field_off + (int64_t)(L + j) * 2 * n;
        const double d = g.d;
        const cd* w = sb < 0 ? ws + lay[conical_w(r)] : nullptr;
        const cd* v = sb < 0 ? ws + lay[conical_v(r)] : nullptr;
        const cd* marks = fields + in.field_off + 4 * (int64_t)(L + 1) * n;
        if (marks[j].x != 0.0) {
            // a Schur-basis layer: p = exp(-k0 z G) a, m = exp(-k0 (d - z) G) b, one triangular exponential per
            // depth and side (field_schur_action, rcwa1d_conical_fields.hpp:106-156)
            const int64_t nn = (int64_t)n * n;
            cd* av = E2 + NC;
            cd* bv = av + n;
            cd* dg = bv + n;
            cd* S1 = dg + n;
            cd* S2 = S1 + nn;
            cd* S3 = S2 + nn;
            cd* S4 = S3 + nn;
            const LayerView lv = layer_at(st, (int)lay[conical_orig(L, j)]);
            const cd* Gm = schur_generator_keep(const_cast<cd*>(ws), lay, in, j, lv);
            int live = 0;
            for (int i = threadIdx.x; i < n; i += NT) {
                av[i] = ap[i] * j0 + ap[n + i] * j1;
                bv[i] = am[i] * j0 + am[n + i] * j1;
                live |= (av[i].x != 0.0 || av[i].y != 0.0 ? 1 : 0) | (bv[i].x != 0.0 || bv[i].y != 0.0 ? 2 : 0);
            }
            int any = 0;
            for (int bit = 1; bit <= 2; bit <<= 1) any |= block_max_int(live & bit);
            int fail = 0;
            for (int col = 0; col < cnt; ++col) {
                const double from_top = fmax(0.0, fmin(d, zs[cols[g.first + col]] - g.z_top));
                for (int side = 0; side < 2; ++side) {
                    const cd* anc = side ? bv : av;
                    cd* dst = (side ? Q : P) + (int64_t)col * n;
                    if (!(any & (side ? 2 : 1))) {
                        for (int i = threadIdx.x; i < n; i += NT) dst[i] = cd{0.0, 0.0};
                        __syncthreads();
                        continue;
                    }
                    const cd f{-k0 * (side ? d - from_top : from_top), 0.0};
                    for (int64_t e = threadIdx.x; e < nn; e += NT) S1[e] = Gm[e] * f;
                    __syncthreads();
                    for (int i = threadIdx.x; i < n; i += NT) dg[i] = S1[i + (int64_t)i * n];
                    __syncthreads();
                    fail |= wg_triangular_exp(n, S1, dg, S2, S3, S4);
                    for (int i = threadIdx.x; i < n; i += NT) {
                        cd acc{0.0, 0.0};
                        for (int k = 0; k < n; ++k) acc = acc + S2[i + (int64_t)k * n] * anc[k];
                        dst[i] = acc;
                    }
                    __syncthreads();
                }
            }
            const double nan = fail ? __builtin_nan("") : 0.0;
            for (int64_t e = threadIdx.x; e < Nc; e += NT) {
                const cd pp = P[e], mm = Q[e];
                P[e] = fail ? cd{nan, nan} : pp + mm;
                Q[e] = fail ? cd{nan, nan} : mm - pp;
            }
        } else
        for (int64_t e = threadIdx.x; e < Nc; e += NT) {
            const int i = (int)(e % n), col = (int)(e / n);
            const double from_top = fmax(0.0, fmin(d, zs[cols[g.first + col]] - g.z_top));
            const double from_bottom = d - from_top;
            const cd a = ap[i] * j0 + ap[n + i] * j1, b = am[i] * j0 + am[n + i] * j1;
            const bool az = a.x == 0.0 && a.y == 0.0, bz = b.x == 0.0 && b.y == 0.0;
            const cd pp = az ? cd{0.0, 0.0} : a * cexp((-k0 * from_top) * q[i]);
            const cd mm = bz ? cd{0.0, 0.0} : b * cexp((-k0 * from_bottom) * q[i]);
            if (sb < 0) {
                P[e] = w[i] * (pp + mm);
                Q[e] = v[i] * (mm - pp);
            } else {
                P[e] = pp + mm;
                Q[e] = mm - pp;
            }
        }
        __syncthreads();
        if (sb >= 0) {
            const ConicalSlots sl = slots_of(const_cast<cd*>(ws) + sb, N);
            wg_gemm<GEMM_SET, false>(n, cnt, n, sl.W, n, P, n, F, n);
            wg_gemm<GEMM_SET, false>(n, cnt, n, sl.V, n, Q, n, G, n);
            Ff = F;
            Gf = G;
            IE = sl.IE;
        } else {
            const LayerView lv = layer_at(st, (int)lay[conical_orig(L, j)]);
            const cd index = lv.homogeneous ? lv.n : cd{lv.vals[0], lv.vals[1]};
            eps_scale = cdiv(cd{1.0, 0.0}, index * index);
        }
    } else {
        const bool top = g.region == -1;
        const int r = top ? 0 : L + 1;
        const cd* w = ws + lay[conical_w(r)];
        const cd* v = ws + lay[conical_v(r)];
        const cd* resp = fields + in.field_off + 4 * (int64_t)L * n + (top ? 0 : 2 * n);
        const cd kz0 = k0 * geo.kzI[M];
        for (int64_t e = threadIdx.x; e < Nc; e += NT) {
            const int i = (int)(e % n), col = (int)(e / n), o = i % N;
            const double z = zs[cols[g.first + col]];
            const cd rsp = resp[i] * j0 + resp[n + i] * j1;
            const bool rz = rsp.x == 0.0 && rsp.y == 0.0;
            const cd zero{0.0, 0.0};
            cd a = zero, b = zero, t = zero;
            if (top) {
                if (o == M) a = (i < N ? j0 : j1) * cexp(cd{-kz0.y, kz0.x} * z);
                if (!rz) {
                    const cd kz = k0 * geo.kzI[o];
                    b = rsp * cexp(cd{kz.y, -kz.x} * z);
                }
            } else if (!rz) {
                const cd kz = k0 * geo.kzII[o];
                t = rsp * cexp(cd{-kz.y, kz.x} * (z - g.z_bottom));
            }
            P[e] = w[i] * ((a + b) + t);
            Q[e] = v[i] * ((b - a) - t);
        }
        __syncthreads();
        const cd nn = top ? cd{st[1], st[2]} : cd{st[3], st[4]};
        eps_scale = cdiv(cd{1.0, 0.0}, nn * nn);
    }
    // rotate (s, t) -> (x, y) per order (rcwa1d_conical_fields.hpp:210-220); Ez before [[eps]]^-1 into E1
    for (int64_t e = threadIdx.x; e < NC; e += NT) {
        const int jj = (int)(e % N), col = (int)(e / N);
        const int iz = cols[g.first + col];
        const double cs = geo.cs[jj].x, sn = geo.sn[jj].x, kx = geo.kx[jj].x;
        const cd Fa = Ff[jj + (int64_t)col * n], Fb = Ff[N + jj + (int64_t)col * n];
        const cd Ga = Gf[jj + (int64_t)col * n], Gb = Gf[N + jj + (int64_t)col * n];
        const cd Ex = -sn * Fa + cs * Fb;
        const cd Ey = cs * Fa + sn * Fb;
        const cd hx = -cs * Ga - sn * Gb;
        const cd hy = -sn * Ga + cs * Gb;
        const cd u = ky * hx - kx * hy;
        out[0 * NZ + jj + (int64_t)iz * N] = Ex;
        out[1 * NZ + jj + (int64_t)iz * N] = Ey;
        out[3 * NZ + jj + (int64_t)iz * N] = cd{hx.y, -hx.x};
        out[4 * NZ + jj + (int64_t)iz * N] = cd{hy.y, -hy.x};
        out[5 * NZ + jj + (int64_t)iz * N] = ky * Ex - kx * Ey;
        E1[e] = cd{-u.y, u.x};
    }
    __syncthreads();
    if (IE) {
        wg_gemm<GEMM_SET, false>(N, cnt, N, IE, N, E1, N, E2, N);
        for (int64_t e = threadIdx.x; e < NC; e += NT)
            out[2 * NZ + (e % N) + (int64_t)cols[g.first + (int)(e / N)] * N] = E2[e];
    } else {
        for (int64_t e = threadIdx.x; e < NC; e += NT)
            out[2 * NZ + (e % N) + (int64_t)cols[g.first + (int)(e / N)] * N] = E1[e] * eps_scale;
    }
}

void launch_fields_conical(hipStream_t s, int n_groups, const Inst1D* inst, const double* stacks,
                           const int64_t* layouts, const cd* ws, const cd* fields, const FieldGroup1D* groups,
                           const int* cols, const double* zs, int nz, cd j0, cd j1, cd* out, cd* scratch) {
    k_fields_conical<<<n_groups, NT, 0, s>>>(inst, stacks, layouts, ws, fields, groups, cols, zs, nz, j0, j1, out, scratch);
}

void launch_conical(hipStream_t s, int n_inst, int n_tasks, const Inst1D* insts, const LayerTask* tasks,
                    const double* stacks, const int64_t* layouts, cd* ws, int* task_info, Out1D out, int max_iter, int min_n, int max_n,
                    int part, int flags, hipEvent_t* ev, const int* eig_order, const OpsArgs& ops) {
    const bool diag = flags & PLAN_DIAGNOSTICS;
    // part 1: shared operators (first chunk), setup, assembly, eigensolves (events ev[0..2]); part 2:
    // modes, S-matrix, efficiencies (events ev[3], ev[4]); the host may run the parts of different chunks
    // on two streams
    QcLaunch a{};
    a.stream = s;
    a.insts = insts;
    a.tasks = tasks;
    a.stacks = stacks;
    a.layouts = layouts;
    a.ws = ws;
    a.task_info = task_info;
    a.max_iter = max_iter;
    a.order = eig_order;
    a.op_descs = ops.descs;
    a.ops = ops.ops;
    if (part & 1) {
        if (ev) hipEventRecord(ev[0], s);
        if (ops.compute && ops.n > 0)
            k_ops_conical<<<ops.n, NT, 0, s>>>(ops.descs, static_cast<const Inst1D*>(ops.insts), stacks, ops.ops,
                                               ops.info);
        k_setup_conical<<<n_inst, NT, 0, s>>>(insts, stacks, layouts, ws, out.geometry);
        if (n_tasks > 0)
            k_assemble_conical<<<n_tasks, NT, 0, s>>>(insts, tasks, stacks, layouts, ws, task_info, out.task_diag,
                                                      ops.descs, ops.ops, ops.info);
        if (ev) hipEventRecord(ev[1], s);
        // each matrix goes to the wave or the workgroup eigensolver by its own order
        if (n_tasks > 0 && min_n <= WAVE_EIG_MAX_N) {
            a.what = 1;
            a.n_items = 2 * n_tasks;
            a.grid = (2 * n_tasks + NWAVE - 1) / NWAVE;
            conical_qc(wave_columns_per_lane(WAVE_EIG_MAX_N), a);
        }
        if (n_tasks > 0 && max_n > WAVE_EIG_MAX_N) {
            a.what = 0;
            a.grid = 2 * n_tasks;
            conical_qc(eig_columns_per_thread(max_n), a);
        }
        if (n_tasks > 0 && diag) {
            k_eigres_conical<<<2 * n_tasks, NT, 0, s>>>(insts, tasks, layouts, ws, task_info, out.item_res);
            k_fold_res<<<(n_tasks + NT - 1) / NT, NT, 0, s>>>(n_tasks, out.item_res, out.task_diag);
        }
        if (ev) hipEventRecord(ev[2], s);
    }
    if (part & 2) {
        if (ev) hipEventRecord(ev[3], s);
        if (n_tasks > 0) {
            a.what = 2;
            a.grid = n_tasks;
            conical_qc(eig_columns_per_thread(2 * max_n), a);
        }
        if (out.fields) smatrix_conical_fields_launch(n_inst, s, insts, stacks, layouts, ws, task_info, out);
        else smatrix_conical_launch(n_inst, s, insts, stacks, layouts, ws, task_info, out);
        if (ev) hipEventRecord(ev[4], s);
    }
}
#endif
