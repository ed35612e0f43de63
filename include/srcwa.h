/* srcwa.h -- C-ABI of the MI355X-native RCWA solver (libsrcwa_amd.so).
 *
 * Plain pointers and sizes only; every entry point is exception-free and returns an int code
 * (SRCWA_OK or a negative SRCWA_E*), with a message for the calling thread in srcwa_last_error().
 * Complex numbers are interleaved (re, im) doubles; matrices are column-major, as the reference's
 * stablercwa::Matrix (include/stablercwa/matrix.hpp:1-5).
 *
 * Flat stack records (doubles):
 *   planar / conical stack (reference rcwa::Stack, rcwa1d.hpp:80-114):
 *     period, nI.re, nI.im, nII.re, nII.im, L, then per layer
 *       d, kind (0 uniform | 1 grating),
 *         uniform: n.re, n.im
 *         grating: K, edges[K+1], (n.re, n.im)[K], fourier_method (0 analytic | 1 FFT), fft_samples
 *   crossed-grating stack (reference rcwa::Stack2D, rcwa2d_types.hpp:39-78):
 *     Lx, Ly, nI, nII.re, nII.im, L, then per layer
 *       d, kind (0 uniform | 1 patterned),
 *         uniform: n.re, n.im
 *         patterned: Ky, y_edges[Ky+1], then per y segment Kx, x_edges[Kx+1], (n.re, n.im)[Kx]
 *
 * Status words per solve: 0 ok, 1 numerical failure (results NaN; the reference's NumericalFailure,
 * rcwa1d.hpp:37-38), 2 device workspace exhausted (a sizing error of this library, never expected),
 * 3 Rayleigh anomaly (crossed gratings, rcwa2d_types.hpp:32-37).
 */
#ifndef SRCWA_H
#define SRCWA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SRCWA_OK = 0,
    SRCWA_EINVAL = -1,        /* invalid argument: the reference throws std::invalid_argument */
    SRCWA_EDEVICE = -2,       /* HIP runtime error */
    SRCWA_ENOMEM = -3,        /* device allocation failed */
    SRCWA_EUNSUPPORTED = -4,  /* option not provided by this build */
    SRCWA_EUNAVAILABLE = -5,  /* a near-field context is unavailable: the reference throws std::logic_error */
    SRCWA_ENUMERICAL = -6,    /* a near-field evaluation is not finite: the reference throws std::runtime_error */
};

typedef struct srcwa_context srcwa_context;
typedef struct srcwa_plan srcwa_plan;

const char* srcwa_last_error(void);
const char* srcwa_version(void);

/* One context per device: a HIP stream and the device it runs on. */
int srcwa_context_create(int device, srcwa_context** out);
void srcwa_context_destroy(srcwa_context* ctx);
int srcwa_context_sync(srcwa_context* ctx);

/* A device memory pool: plans created on a context with a pool take all their device buffers from
 * it (first fit, 256-byte blocks) instead of hipMalloc, and fail with SRCWA_ENOMEM when it is full.
 * Several contexts may share one pool.  Destroy plans before the pool. */
typedef struct srcwa_pool srcwa_pool;
int srcwa_pool_create(int device, int64_t bytes, srcwa_pool** out);
int srcwa_pool_destroy(srcwa_pool* pool);
int64_t srcwa_pool_free_bytes(srcwa_pool* pool);
int srcwa_context_set_pool(srcwa_context* ctx, srcwa_pool* pool);

/* Flags of the plans created on a context from now on (they are fixed at plan creation).
 * SRCWA_PLAN_DIAGNOSTICS: the reference's Options::diagnostics (rcwa1d.hpp:120, 141-163): every layer
 *   eigensolve's eigen-pair residual max_j ||A v_j - l_j v_j|| / (||A||_F ||v_j||) (eig.hpp:38-69) and every
 *   dense solve's backward error ||A X - B||_F / (||A||_F ||X||_F + ||B||_F) (lu.hpp:452-459), reduced per
 *   solve by srcwa_plan_diagnostics; the plan holds copies of the layer matrices (more device memory)
 *   and each dense solve does one more product.
 * SRCWA_PLAN_MIXED: the S-matrix recursion's dense solves (the star products' (I - R r)^{-1} T and the
 *   interface, top-block and downward-sweep solves) are factored in complex FP32, with the trailing update
 *   on FP32 MFMA, and refined to FP64 accuracy by FP64 residuals (LAPACK zcgesv's per-column stopping
 *   test, at most 6 corrections); a solve whose corrections stall falls back to the FP64 LU.  Eigensolves,
 *   operator assembly and the efficiency sums stay FP64.  No counterpart in the reference (its LU is FP64
 *   only, lu.hpp:267-443); srcwa_plan_refinement reports what each solve needed.
 * SRCWA_PLAN_POINT_OPS: every grating layer of every point computes its own wavelength-invariant
 *   operators, as the reference's solve does (rcwa1d.hpp:432-468, rcwa1d_conical.hpp:302-336,
 *   rcwa2d_operators.hpp:349-429).  Without it a plan computes them once per execution for all of its
 *   points that share them (the same layer record apart from the thickness, period or cell, truncation and,
 *   planar, polarization): a sweep over wavelength or angle, or over the thickness of a layer, factors
 *   each layer's operators once.  Results are the same bits either way; diagnostics plans always use the
 *   per-point form.
 * SRCWA_PLAN_FIELDS: planar (TE/TM) plans keep each solve's near-field capture, as rcwa::solve_with_fields
 *   does (rcwa1d.hpp:480-489, 1006-1023; rcwa1d_fields.hpp:55-104): the layer amplitudes a+ (layer top) and
 *   a- (layer bottom), beside the modes and R, T the solve leaves on the device; srcwa_plan1d_field_at and
 *   srcwa_plan1d_field_map evaluate them.  Conical plans keep theirs the same way (rcwa1d_conical.hpp:859-970;
 *   srcwa_plan_conical_field_at / _field_map).  Such a plan must fit one workspace chunk.  Crossed-grating
 *   plans refuse the flag (SRCWA_EUNSUPPORTED). */
enum { SRCWA_PLAN_DIAGNOSTICS = 1, SRCWA_PLAN_MIXED = 2, SRCWA_PLAN_POINT_OPS = 4, SRCWA_PLAN_FIELDS = 8 };
int srcwa_context_set_plan_flags(srcwa_context* ctx, int flags);

/* ---------------------------------------------------------------------------------------------
 * Batched planar (TE/TM) solves.  Replaces rcwa::solve / rcwa::sweep (rcwa1d.hpp:1077-1091): each
 * problem is (stack, wavelength, theta_deg, M, pol) with pol 0 = TE, 1 = TM, and yields N = 2M+1
 * orders of R, T (complex), DEr, DEt (real), a status word and eig_info (Result, rcwa1d.hpp:145-164).
 * Results of problem i start at order offset sum_{k<i} (2 M_k + 1).
 *
 * stacks: n_stacks flat records concatenated; stack_offsets[s] = start (in doubles) of record s.
 * max_workspace_bytes bounds the device workspace of one launch chunk (0 = 64 GiB); problems are
 * admitted into chunks in order until the bound is reached.  With several chunks, execution runs on
 * the context stream and a second stream of the context: the eigensolves of chunk c+1 overlap the
 * S-matrix stage of chunk c, and two workspace regions alternate between chunks (device workspace =
 * 2 x the largest chunk's).  Results do not depend on the chunking.
 */
int srcwa_plan1d_create(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                        int n_problems, const int* stack_index, const double* wavelength, const double* theta_deg,
                        const int* M, const int* pol, int64_t max_workspace_bytes, srcwa_plan** out);

/* Batched conical solves.  Replaces rcwa::solve_conical / sweep_conical (rcwa1d_conical.hpp:985-999).
 * Per problem N = 2M+1 orders; R and T receive the Jones matrices Jr, Jt as N x 4 complex
 * (ss, sp, ps, pp per order) and DEr, DEt the N x 4 channel efficiencies flux * |J|^2
 * (ConicalResult, rcwa1d_conical.hpp:105-183).  Result offsets are 4 * the planar order offsets.
 * Degenerate layers (eta = min|lambda| / |ky|^1.5 < 1e-2) take the Schur-basis propagation on the GPU
 * (rcwa1d_conical.hpp:413-502, 797-835). */
int srcwa_plan_conical_create(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                              int n_problems, const int* stack_index, const double* wavelength,
                              const double* theta_deg, const double* phi_deg, const int* M,
                              int64_t max_workspace_bytes, srcwa_plan** out);

/* Batched crossed-grating (2-D) solves.  Replaces rcwa::solve2d (rcwa2d.hpp:1042-1411) over many
 * (stack, wavelength, theta, phi, M1, M2, jones, factorization) problems.
 *
 * stacks: flat crossed-grating records (Lx, Ly, nI, nII.re, nII.im, L, then per layer d, kind,
 *   uniform: n.re, n.im | patterned: Ky, y_edges[Ky+1], per y segment Kx, x_edges[Kx+1], (n.re, n.im)[Kx];
 *   reference Stack2D, rcwa2d_types.hpp:39-63).  Patterned layers hold at most 256 y segments.
 * jones: 4 doubles per problem, [Es.re, Es.im, Ep.re, Ep.im] (rcwa2d_types.hpp:97-107).
 * factorization: NULL or one int per problem, 0 = Li (default), 1 = Laurent (Factorization2D).
 * Per problem N = (2 M1 + 1)(2 M2 + 1) orders, harmonic i = (m + M1)(2 M2 + 1) + (n + M2).
 * Results (srcwa_plan_results): R = [Rs, Rp, Ex_r, Ey_r], T = [Ts, Tp, Ex_t, Ey_t] (N complex each) at
 * 4 x the order offset;
 * DEr, DEt (N) at the order offset; status 0 ok, 1 numerical failure, 3 Rayleigh anomaly (Status2D).
 * Invalid inputs return SRCWA_EINVAL with the reference's std::invalid_argument message
 * (check_core_inputs / check_layer_core / check_patterned, rcwa2d.hpp:953-1026). */
int srcwa_plan2d_create(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                        int n_problems, const int* stack_index, const double* wavelength, const double* theta_deg,
                        const double* phi_deg, const int* M1, const int* M2, const double* jones,
                        const int* factorization, int64_t max_workspace_bytes, srcwa_plan** out);

/* Device bytes a plan with these arguments would hold, without allocating anything; the inputs are
 * validated exactly as by the matching *_create call (same error codes and messages). */
int srcwa_plan1d_bytes(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                       int n_problems, const int* stack_index, const double* wavelength, const double* theta_deg,
                       const int* M, const int* pol, int64_t max_workspace_bytes, int64_t* bytes);
int srcwa_plan_conical_bytes(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                             int n_problems, const int* stack_index, const double* wavelength,
                             const double* theta_deg, const double* phi_deg, const int* M,
                             int64_t max_workspace_bytes, int64_t* bytes);
int srcwa_plan2d_bytes(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                       int n_problems, const int* stack_index, const double* wavelength, const double* theta_deg,
                       const double* phi_deg, const int* M1, const int* M2, const double* jones,
                       const int* factorization, int64_t max_workspace_bytes, int64_t* bytes);

/* Capture the plan's launch sequence into a HIP graph (enable = 1) that srcwa_plan_execute then
 * replays with one hipGraphLaunch; enable = 0 returns to launching kernel by kernel.  One-chunk plans
 * only (SRCWA_EUNSUPPORTED otherwise).  A replayed plan reports only the total in srcwa_plan_stage_ms
 * (the stage entries are -1). */
int srcwa_plan_set_graph(srcwa_plan* plan, int enable);
/* Iteration limit of every layer eigensolve of the plan: QR sweeps per eigenvalue before the solve
 * reports eig_info = the index (1-based) of the eigenvalue that did not converge, with NaN results
 * (EigOptions::max_iter, eig_types.hpp; default 30, as the reference).  A captured graph is captured
 * again.  Lets a caller reproduce the reference's non-convergence contract
 * (tests/test_rcwa1d.cpp:1169-1181). */
int srcwa_plan_set_max_iter(srcwa_plan* plan, int max_iter);
/* Launch all chunks of the plan on the context stream (asynchronous). */
int srcwa_plan_execute(srcwa_plan* plan);
/* 1 when the last execute has finished, 0 while it runs, a negative code on a device error. */
int srcwa_plan_poll(srcwa_plan* plan);
/* Wait for the last execute. */
int srcwa_plan_sync(srcwa_plan* plan);
/* Copy results to host buffers (implies sync).  Any pointer may be NULL. */
int srcwa_plan_results(srcwa_plan* plan, double* R, double* T, double* DEr, double* DEt, int* status, int* eig_info);
/* Per-solve diagnostics of the last execute, 5 doubles per problem (Result::max_eig_residual,
 * max_solve_backward_error, n_eig, n_dense_solve; rcwa1d.hpp:159-162, ConicalResult, Result2D; and
 * ConicalResult::n_schur, rcwa1d_conical.hpp:175):
 * [0] largest eigen-pair residual of the solve's eigensolves (NaN when one did not converge; 0 without
 * SRCWA_PLAN_DIAGNOSTICS), [1] largest backward error of its dense solves (0 without the flag, and for
 * crossed gratings, whose Result2D has no such field), [2] dense eigensolves, [3] dense linear solves,
 * counted as the reference counts them (a conical layer in the Schur basis adds one eigensolve), [4] conical
 * layers propagated in the Schur basis (0 for planar and crossed plans). */
int srcwa_plan_diagnostics(srcwa_plan* plan, double* out);
/* The mixed-precision record of the last execute, 3 doubles per problem (zeros for plans without
 * SRCWA_PLAN_MIXED): [0] dense solves that fell back to the FP64 LU, [1] FP32 correction steps over all of
 * its refined solves, [2] the largest final max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf) of its
 * refined solves (0 when every solve fell back).  No counterpart in the reference. */
int srcwa_plan_refinement(srcwa_plan* plan, double* out);
/* Per-order geometry of the last execute of a conical or crossed plan (8 doubles per order, at 8 x the
 * order offset).  Conical: kx, kzI (re, im), kzII (re, im), flux_r, flux_t, phi_m in degrees
 * (ConicalResult::kx, kzI, kzII, flux_r, flux_t, phi_m_deg, rcwa1d_conical.hpp:105-112).  Crossed: Kx, Ky,
 * kzI (re, im), kzII (re, im), 0, 0 (Result2D::Kx, Ky, kzI, kzII, rcwa2d_types.hpp:119-121). */
int srcwa_plan_geometry(srcwa_plan* plan, double* out);
/* Total result orders (sum of N over problems), device bytes held, number of launch chunks. */
int64_t srcwa_plan_orders(const srcwa_plan* plan);

/* Near fields of problem `problem` of an executed planar plan created with SRCWA_PLAN_FIELDS.
 * srcwa_plan1d_field_available: 1 when the problem's capture is usable (PlanarFieldContext::available(),
 *   rcwa1d_fields.hpp:33; the solve succeeded and every captured mode and amplitude is finite), else 0.
 * srcwa_plan1d_field_at: the harmonic fields at nz depths z (field_at, rcwa1d_fields.hpp:193-198): out holds
 *   6 components (Ex, Ey, Ez, Hx, Hy, Hz; H is Z0 H) of N x nz complex, column-major (order i, depth iz at
 *   i + iz N), component c from c N nz; depths keep their order and repeats (HarmonicFields,
 *   rcwa_fields.hpp:108-128).  z < 0 is the incidence side, z past the stack the exit side; an interface
 *   belongs to the layer above it.
 * srcwa_plan1d_field_map: the same fields on the x grid x_k = k period / nx, k < nx, at height y
 *   (field_map, rcwa1d_fields.hpp:200-206; reconstruct_field_grid, rcwa_fields.hpp:326-355): 6 components of
 *   nx x nz complex, (x_k, z_iz) at k + iz nx.
 * Errors: SRCWA_EINVAL (non-finite depth, nx <= 0, non-finite y, a plan without fields or not executed),
 * SRCWA_EUNAVAILABLE (the problem's capture is unusable), SRCWA_ENUMERICAL (a non-finite field value). */
int srcwa_plan1d_field_available(srcwa_plan* plan, int problem, int* available);
int srcwa_plan1d_field_at(srcwa_plan* plan, int problem, int nz, const double* zs, double* out);
int srcwa_plan1d_field_map(srcwa_plan* plan, int problem, int nz, const double* zs, int nx, double y, double* out);
/* Per layer of an available planar capture: its modes with q exactly zero, which propagate in the linear
 * finite-zero basis (FieldLayerInfo::finite_zero_mode_count, rcwa1d_fields.hpp:94-103); one int per layer. */
int srcwa_plan1d_field_layers(srcwa_plan* plan, int problem, int* zero_modes);

/* Near fields of problem `problem` of an executed conical plan created with SRCWA_PLAN_FIELDS, for the
 * incident Jones vector jones = [Es.re, Es.im, Ep.re, Ep.im] (normalized first, rcwa_fields.hpp:218-230): as
 * srcwa_plan1d_field_at / _field_map (ConicalFieldContext, field_at, field_map,
 * rcwa1d_conical_fields.hpp:14-262), with the grid's transverse factor exp(-i (kxm x + ky y)).  Layers of
 * zero thickness hold no fields (the depth groups skip them); degenerate layers are sampled in their Schur
 * basis (exp(-k0 z G) per depth).  srcwa_plan1d_field_available serves conical plans too.  A Jones vector
 * that is not finite or is zero: SRCWA_EINVAL. */
int srcwa_plan_conical_field_at(srcwa_plan* plan, int problem, int nz, const double* zs, const double* jones,
                                double* out);
/* Per layer of positive thickness (in stack order) of an available conical capture: 1 when the layer
 * propagates in the Schur basis (FieldPropagation::Schur of the field metadata, rcwa1d_conical_fields.hpp:89-96),
 * else 0; schur holds one int per such layer. */
int srcwa_plan_conical_field_layers(srcwa_plan* plan, int problem, int* schur);
int srcwa_plan_conical_field_map(srcwa_plan* plan, int problem, int nz, const double* zs, int nx, double y,
                                 const double* jones, double* out);
int64_t srcwa_plan_device_bytes(const srcwa_plan* plan);
int srcwa_plan_chunks(const srcwa_plan* plan);
/* Shared operator sets of the plan (0 with SRCWA_PLAN_POINT_OPS, with diagnostics, or when no two points
 * share a layer's operators).  No counterpart in the reference. */
int srcwa_plan_operator_sets(const srcwa_plan* plan);
/* Per-stage GPU milliseconds of the last execute, from HIP events on the plan's stream:
 * [0] setup + assembly, [1] eigensolves, [2] modes + S-matrix + efficiencies (each summed over
 * chunks), [3] wall time from the first chunk's start to the last chunk's end. */
int srcwa_plan_stage_ms(const srcwa_plan* plan, double* ms4);
/* Algorithmic FP64 flops of one execute of the plan, by the accounting convention
 * (complex GEMM m k n: 8mkn; LU n: 8n^3/3; triangular solves with k RHS: 8 n^2 k; eig with vectors: 100 n^3):
 * [0] eigensolves, [1] everything else. */
int srcwa_plan_flops(const srcwa_plan* plan, double* flops2);
void srcwa_plan_destroy(srcwa_plan* plan);

/* One-shot convenience: create, execute, fetch, destroy. */
int srcwa_solve1d_batch(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                        int n_problems, const int* stack_index, const double* wavelength, const double* theta_deg,
                        const int* M, const int* pol, double* R, double* T, double* DEr, double* DEt, int* status,
                        int* eig_info);

/* ---------------------------------------------------------------------------------------------
 * Batched dense kernels in the shapes of the reference's hooks (Options::custom_eig / custom_solve /
 * custom_multiply, rcwa1d.hpp:122-136; cuda_backend.hpp:387-512 is the reference's NVIDIA wiring).
 * Host buffers, items packed back to back, one workgroup per item.
 *   zgeev : A_i n_i x n_i -> w_i (n_i), V_i (n_i x n_i, unit 2-norm columns), info_i (EigResult::info)
 *   zgesv : A_i n_i x n_i, B_i n_i x nrhs_i -> X_i; info_i = 1 on an exact zero pivot
 *   zgemm : C_i (m_i x n_i) = A_i (m_i x k_i) B_i (k_i x n_i)
 */
int srcwa_zgeev_batched(srcwa_context* ctx, int count, const int* n, const double* A, double* w, double* V, int* info);
int srcwa_zgesv_batched(srcwa_context* ctx, int count, const int* n, const int* nrhs, const double* A, const double* B,
                        double* X, int* info);
/* srcwa_zgesv_batched in the mixed-precision form of SRCWA_PLAN_MIXED: each system factored in complex FP32
 * (trailing updates on FP32 MFMA) and refined to FP64 accuracy by FP64 residuals, falling back to the FP64
 * LU when the corrections stall.  refine: 3 doubles per item -- fell back to FP64 (0 / 1), FP32 correction
 * steps, final max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf) (0 after a fallback).  No counterpart in
 * the reference (its LU is FP64 only, lu.hpp:267-443); a custom_solve hook may bind it. */
int srcwa_zgesv_batched_mixed(srcwa_context* ctx, int count, const int* n, const int* nrhs, const double* A,
                              const double* B, double* X, int* info, double* refine);
/* srcwa_zgeev_batched with the QR iteration limit max_iter per eigenvalue (EigOptions::max_iter,
 * eig_types.hpp:29; 30 in srcwa_zgeev_batched): info_i = k > 0 when eigenvalue k (1-based) did not converge,
 * and V_i is then not meaningful.  The custom_eig adapter of include/stablercwa_amd/hooks.hpp binds it. */
int srcwa_zgeev_batched_iter(srcwa_context* ctx, int count, const int* n, const double* A, double* w, double* V,
                             int* info, int max_iter);
int srcwa_zgemm_batched(srcwa_context* ctx, int count, const int* m, const int* n, const int* k, const double* A,
                        const double* B, double* C);

/* srcwa_zgeev_batched without any device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  A, w, V and info are DEVICE pointers the caller owns, items packed back to back
 * as in srcwa_zgeev_batched (A_i, V_i: n_i x n_i column-major complex; w_i: n_i complex; info: count
 * ints); n is a host array of the orders.  A is read (V may equal A: then it is decomposed in place).
 * The scratch (item descriptors and each item's Z, X, ort, tmp: 16 (2 n_i^2 + 2 n_i) bytes, rounded up)
 * comes from a 64 MiB device array the library holds on every GPU it runs on; a batch that needs more
 * returns SRCWA_ENOMEM and writes nothing.  Nothing in the scratch has to start at zero.  Synchronous
 * on the context's stream.  Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (from several threads or contexts) race and must be serialised by the caller. */
int srcwa_zgeev_batched_noalloc(srcwa_context* ctx, int count, const int* n, const double* A, double* w, double* V,
                                int* info);

/* ---------------------------------------------------------------------------------------------
 * Request server with continuous batching (no counterpart in the reference, whose solves are
 * synchronous calls; it serves the reference's solve/sweep requests to many callers at once).
 *
 * A worker thread keeps up to `slots` batches on the GPU, each on its own HIP stream, with all device
 * memory drawn from one pool of hbm_bytes.  Requests that arrive while batches run are admitted as
 * soon as a slot and memory free up; a batch holds requests of one kind in arrival order, up to
 * max_batch_points points.  Every result is bitwise the same as solving the request alone.
 *
 * Result layout of a request with P points and N orders per point (N = 2M+1, or (2M1+1)(2M2+1)):
 *   planar : R, T  P*N complex;            DEr, DEt P*N
 *   conical: R, T  P*N*4 complex (Jones);  DEr, DEt P*N*4
 *   crossed: R, T  P*4N complex ([Rs, Rp, Ex, Ey] per point); DEr, DEt P*N
 *   status, eig_info: P ints (as srcwa_plan_results).
 */
typedef struct srcwa_server srcwa_server;
typedef struct srcwa_request {
    int kind;                  /* 1 planar, 2 conical, 3 crossed grating */
    int n_points;              /* points of the sweep (>= 1) */
    const double* stack;       /* one flat stack record (planar layout for kinds 1, 2; crossed for 3) */
    int64_t stack_len;
    const double* wavelength;  /* n_points each */
    const double* theta_deg;
    const double* phi_deg;     /* kinds 2, 3 */
    int M;                     /* planar / conical M; crossed M1 */
    int M2;                    /* crossed */
    int pol;                   /* planar: 0 TE, 1 TM */
    int factorization;         /* crossed: 0 Li, 1 Laurent */
    const double* jones;       /* crossed: Es.re, Es.im, Ep.re, Ep.im */
} srcwa_request;

int srcwa_server_create(int device, int64_t hbm_bytes, int slots, int max_batch_points, srcwa_server** out);
/* Validates the request (same errors as the plan calls) and queues it; inputs are copied. */
int srcwa_server_submit(srcwa_server* server, const srcwa_request* request, int64_t* id);
/* Waits up to timeout_s seconds (negative: no limit); *done = 1 once the request has finished. */
int srcwa_server_wait(srcwa_server* server, int64_t id, double timeout_s, int* done);
/* Copies a finished request's results (any pointer may be NULL); times[0] = submit-to-finish seconds,
 * times[1] = seconds queued before its batch started. */
int srcwa_server_result(srcwa_server* server, int64_t id, double* R, double* T, double* DEr, double* DEt,
                        int* status, int* eig_info, double* times);
int srcwa_server_release(srcwa_server* server, int64_t id);
/* batches launched, most batches in flight at once, requests finished, points finished */
int srcwa_server_stats(srcwa_server* server, int64_t* stats4);
const char* srcwa_server_last_error(void);
void srcwa_server_destroy(srcwa_server* server);

#ifdef __cplusplus
}
#endif

#endif /* SRCWA_H */
