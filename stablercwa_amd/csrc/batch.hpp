// Device-side descriptors shared by the host runtime and the kernels.
#pragma once

#include <cstdint>

// Largest matrix order the eigensolver's LDS rotation table holds.
constexpr int EIG_MAX_N_HOST = 1024;

// Status word of one instance.
enum : int {
    ST_OK = 0,
    ST_NUMERICAL = 1,      // singular system, non-finite value or unconverged eigensolve: results are NaN
    ST_WORKSPACE = 2,      // internal: the instance's device workspace was too small (never expected)
    ST_RAYLEIGH = 3,       // crossed gratings: an order is exactly grazing in a homogeneous medium
    ST_SCHUR = 4,          // conical: a degenerate layer needs Schur-basis propagation (not provided)
};

// Plan flags (srcwa_context_set_plan_flags), carried per instance.
enum : int {
    PLAN_DIAGNOSTICS = 1,      // eigen-pair residuals and solve backward errors (Options::diagnostics)
    PLAN_MIXED = 2,            // S-matrix dense solves by FP32 LU + FP64 iterative refinement
    PLAN_POINT_OPS = 4,        // every layer task computes its own wavelength-invariant operators
    PLAN_FIELDS = 8,           // planar: keep the near-field capture (layer amplitudes a+, a-) of every solve
};

// Per-layer-task diagnostics, reduced per instance by the S-matrix kernels.
struct TaskDiag {
    double eig_res;            // largest eigen-pair residual of the task's eigensolves (NaN: no vectors)
    double berr;               // largest backward error of the task's dense solves
    int n_solve;               // dense linear solves of the task (assembly)
    int n_eig;                 // dense eigensolves of the task
};

// Per-instance diagnostics record: max_eig_residual, max_solve_backward_error, n_eig, n_dense_solve,
// n_schur (srcwa_plan_diagnostics), then the mixed-precision record (srcwa_plan_refinement): refined
// solves that fell back to FP64, FP32 correction steps, largest final refined residual.
constexpr int DIAG_WORDS = 8;
constexpr int DIAG_REPORT_WORDS = 5;

// One planar (TE/TM) or conical instance.
struct Inst1D {
    int M, N, pol, L;          // pol: 0 TE, 1 TM, 2 conical
    int first_task, n_tasks;   // layer tasks (grating layers) of this instance
    int flags;                 // PLAN_* flags
    double wavelength, theta_deg, phi_deg;
    int64_t stack_off;         // offset of the flat stack record (doubles)
    int64_t layout_off;        // offset of the layout table (int64)
    int64_t out_off;           // offset of this instance's orders in the result arrays
    int64_t ws_lo, ws_hi;      // bump region of the S-matrix recursion (cd units)
    int64_t field_off;         // PLAN_FIELDS: offset of this instance's a+ (L x N) then a- (L x N), cd units
};

// Samples of an FFT layer (resolve_fft_samples, fourier.hpp:26-44, for inputs the host has validated):
// fft_samples itself, or the smallest power of two >= max(8192, 64 (2M + 1)).
__host__ __device__ inline int fft_samples_for(int M, int samples) {
    if (samples > 0) return samples;
    const long long target = 64LL * (2LL * M + 1) > 8192LL ? 64LL * (2LL * M + 1) : 8192LL;
    int n = 1;
    while (n < target) n *= 2;
    return n;
}

// Scratch of one mixed-precision solve (wg_refine_solve) in complex-double words, as carved up at its
// start: the FP32 factors, x, r, the FP32 right-hand side, the pivots and the gather list.
__host__ __device__ inline int64_t refine_words(int64_t n, int64_t nrhs) {
    return ((n * n + 1) / 2 + 16) + 2 * (n * nrhs + 16) + ((n * nrhs + 1) / 2 + 16) + 2 * (n / 4 + 16);
}
// Bump words a mixed-precision dense solve of order n with up to nrhs right-hand sides takes beyond what
// the FP64 solve takes (the FP64 factors, which the fallback uses, are held by both forms).
__host__ __device__ inline int64_t mixed_solve_words(int64_t n, int64_t nrhs) {
    return refine_words(n, nrhs) + 2 * n * nrhs + 64;
}

// One layer task: a grating layer of an instance that needs an eigensolve.  op: the task's shared
// operator set (below), or -1 when the task computes its layer's operators itself.
struct LayerTask {
    int inst, layer;
    int op = -1;
};

// A shared operator set: the wavelength-invariant operators of one grating layer (the Toeplitz matrices
// of eps and 1/eps and their LU factors), computed once per plan execution for every task whose layer
// record, period, truncation order and polarization are the same.  inst (a plan-wide instance index) and
// layer name one such task; off is the set's cd offset in the plan's operator buffer.
struct LayerOps {
    int inst, layer;
    int64_t off;
};

// A plan's shared operator sets as the launchers take them: the descriptors, the operator buffer, each
// set's failure word (nonzero: singular or invalid), the plan-wide instance array the descriptors index,
// and whether this launch computes the sets (the first chunk of an execution) before its assembly.
struct cd;
struct OpsArgs {
    const LayerOps* descs = nullptr;
    cd* ops = nullptr;
    int* info = nullptr;
    int n = 0;
    const void* insts = nullptr;
    bool compute = false;
};

// Planar operator set of order n (cd words, each part 16-aligned): TE: [[eps]]; TM: the LU factors of
// [[eps]], [[1/eps]] and the LU factors of [[1/eps]], then the two pivot vectors.  Both: then the 4M+1
// Fourier coefficients (2n - 1) and, for an FFT layer, its sample scratch.
__host__ __device__ inline int64_t ops_align16(int64_t x) { return (x + 15) & ~int64_t(15); }
__host__ __device__ inline int64_t planar_ops_matrices(int pol) { return pol == 0 ? 1 : 3; }
__host__ __device__ inline int64_t planar_ops_words(int64_t n, int pol) {
    return planar_ops_matrices(pol) * ops_align16(n * n) + (pol == 0 ? 0 : 2 * ops_align16(n)) + ops_align16(2 * n);
}

// Conical operator set of order n: [[eps]], [[eps]]^-1, [[1/eps]]^-1, an LU scratch matrix, the pivots,
// the 4M+1 Fourier coefficients and, for an FFT layer, its sample scratch.
__host__ __device__ inline int64_t conical_ops_words(int64_t n) {
    return 4 * ops_align16(n * n) + ops_align16(n) + ops_align16(2 * n);
}

// Crossed-grating operator set of N harmonics: Eps, Cx, Cy, the LU factors of Eps, Eps^-1, the pivots and
// the operator scratch (scratch words, as the per-layer slot's).
__host__ __device__ inline int64_t crossed_ops_words(int64_t N, int64_t scratch) {
    return 5 * ops_align16(N * N) + ops_align16(N) + ops_align16(scratch);
}

// Layout table of a planar instance (int64 cd offsets into the workspace):
//   [0]                      kx (N)
//   [1 + 2r], [2 + 2r]       q, p of region r = 0 .. L+1 (N each)
//   [1 + 2(L+2) + j]         base of layer j's slots, or -1 for a uniform film:
//                            S0..S3 (N x N each), then lambda, ort, tmp (N each), pivots (N ints)
constexpr int LAYOUT_KX = 0;
__host__ __device__ inline int layout_q(int r) { return 1 + 2 * r; }
__host__ __device__ inline int layout_p(int r) { return 2 + 2 * r; }
__host__ __device__ inline int layout_layer(int L, int j) { return 1 + 2 * (L + 2) + j; }
__host__ __device__ inline int layout_size(int L) { return 1 + 2 * (L + 2) + L; }
// slots of a grating layer of order n
__host__ __device__ inline int64_t slot_words(int64_t n) { return 4 * n * n + 4 * n; }

// Layout table of a conical instance (L = layers with d > 0):
//   [0]                          geometry: kx, cos phi_m, sin phi_m, kzI, kzII, flux (N each), ky (1)
//   [1 + 3r], [2 + 3r], [3 + 3r] q, w, v of region r = 0 .. L+1 (2N each; w, v = -1 for grating layers)
//   [1 + 3(L+2) + j]             slot base of layer j (-1 for a uniform layer)
//   [1 + 3(L+2) + L + j]         index of layer j in the stack record
__host__ __device__ inline int conical_q(int r) { return 1 + 3 * r; }
__host__ __device__ inline int conical_w(int r) { return 2 + 3 * r; }
__host__ __device__ inline int conical_v(int r) { return 3 + 3 * r; }
__host__ __device__ inline int conical_slot(int L, int j) { return 1 + 3 * (L + 2) + j; }
__host__ __device__ inline int conical_orig(int L, int j) { return 1 + 3 * (L + 2) + L + j; }
__host__ __device__ inline int conical_layout_size(int L) { return 1 + 3 * (L + 2) + 2 * L; }
__host__ __device__ inline int64_t conical_slot_words(int64_t N) { return 16 * N * N + 7 * N; }

// Layout table of a crossed-grating instance (N harmonics, n = 2N modes):
//   [0]                      geometry: Kx, Ky (real, N each), kzI, kzII (N each), Rayleigh flag (1)
//   [1 + 2r], [2 + 2r]       q (2N) and 2x2 block-diagonal V (a11, a12, a21, a22: 4N) of region r;
//                            the V entry is -1 for patterned layers
//   [1 + 2(L+2) + j]         slot base of patterned layer j (-1 for a uniform layer)
//   [1 + 2(L+2) + L + j]     offset (doubles, in the stack buffer) of layer j's x-cut record [Kc, cuts]
//   [1 + 2(L+2) + 2L + j]    size (cd) of layer j's big scratch region (H, Z, X / operator scratch)
// Slots of a patterned layer: EPS, CX, CY (N x N), BIG, then lambda, ort, tmp (2N each), pivots (N).
__host__ __device__ inline int x2_q(int r) { return 1 + 2 * r; }
__host__ __device__ inline int x2_v(int r) { return 2 + 2 * r; }
__host__ __device__ inline int x2_slot(int L, int j) { return 1 + 2 * (L + 2) + j; }
__host__ __device__ inline int x2_cuts(int L, int j) { return 1 + 2 * (L + 2) + L + j; }
__host__ __device__ inline int x2_big(int L, int j) { return 1 + 2 * (L + 2) + 2 * L + j; }
__host__ __device__ inline int x2_layout_size(int L) { return 1 + 2 * (L + 2) + 3 * L; }

// One crossed-grating instance.
struct Inst2D {
    int M1, M2, N, L;
    int first_task, n_tasks;
    int flags;                 // PLAN_* flags
    int laurent;               // 0: Li factorisation, 1: Laurent rule
    double wavelength, theta_deg, phi_deg;
    double jones[4];           // Es.re, Es.im, Ep.re, Ep.im
    int64_t stack_off, layout_off, out_off, ws_lo, ws_hi;
};

// Per-instance results.
struct Out1D {
    double* R;      // complex, N per instance (planar) / 4N Jones (conical)
    double* T;
    double* DEr;    // N (planar) / 4N (conical)
    double* DEt;
    int* status;
    int* eig_info;
    double* geometry;   // conical / crossed: GEO_WORDS doubles per order (null for planar plans)
    double* diag;       // DIAG_WORDS per instance
    TaskDiag* task_diag;   // per layer task of the launch chunk
    double* item_res;      // conical: the eigen residual of each of a task's two eigensolves (2 per task)
    double* fields;        // planar with PLAN_FIELDS: complex layer amplitudes (Inst1D::field_off), else null
    int* field_ok;         // planar with PLAN_FIELDS: 1 when the instance's capture is whole and finite
};

// One depth group of a planar near-field sample (rcwa_fields.hpp:253-279): the depths of one region.
struct FieldGroup1D {
    int region;           // -1 incidence side, -2 exit side, j >= 0 layer j
    int first, count;     // its depths: zs[cols[first .. first + count)]
    int tm_grating;       // TM in a grating layer: Ez needs a [[eps]] solve
    double z_top, d;      // layer groups: the layer's top and thickness
    double z_bottom;      // the stack's bottom (exit-side phase reference)
    int64_t scratch;      // offset of the group's scratch (cd units)
};

// Per-order geometry written by the conical and crossed setup kernels:
//   conical: kx, kzI.re, kzI.im, kzII.re, kzII.im, flux_r, flux_t, phi_m (degrees)
//   crossed: Kx, Ky, kzI.re, kzI.im, kzII.re, kzII.im, 0, 0
constexpr int GEO_WORDS = 8;

// Batched dense kernels that mirror the reference's hook shapes (custom_eig / custom_solve /
// custom_multiply, rcwa1d.hpp:122-136): one item per workgroup, column-major complex.
struct DenseItem {
    int n, nrhs, k;            // zgemm: C(n x nrhs) = A(n x k) B(k x nrhs); zgesv: A n x n, B n x nrhs
    int64_t a, b, c, w;        // cd offsets into one arena
    int64_t scratch;           // cd offset of per-item scratch
};
