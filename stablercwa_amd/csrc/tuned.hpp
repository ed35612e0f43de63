// Compile-time tuning constants of the eigen and S-matrix kernels (override with -D at build time).  The
// values here are the ones measured best on MI355X: ring depth and eigen occupancy by
// stablercwa_amd/tune.py, the queue depth and the conical / crossed / S-matrix occupancies by A/B builds
// of the same constants on the benchmark batches.
#pragma once

#ifndef SRCWA_EIG_RING
#define SRCWA_EIG_RING 2        // rows each column prefetches ahead in the QR sweep
#endif
// minimum waves per SIMD each family's workgroup eigen kernel is built for (0: the compiler's choice)
#ifndef SRCWA_EIG_OCC_PLANAR
#define SRCWA_EIG_OCC_PLANAR 3
#endif
#ifndef SRCWA_EIG_OCC_CONICAL
#define SRCWA_EIG_OCC_CONICAL 2
#endif
#ifndef SRCWA_EIG_OCC_CROSSED
#define SRCWA_EIG_OCC_CROSSED 2
#endif
#ifndef SRCWA_WAVE_EIG_MAX_N
#define SRCWA_WAVE_EIG_MAX_N 64 // orders up to which one wave solves a matrix
#endif

// SRCWA_EIG_OCC (one value for every family) is what the tuner's variant builds set
#ifdef SRCWA_EIG_OCC
#undef SRCWA_EIG_OCC_PLANAR
#undef SRCWA_EIG_OCC_CONICAL
#undef SRCWA_EIG_OCC_CROSSED
#define SRCWA_EIG_OCC_PLANAR SRCWA_EIG_OCC
#define SRCWA_EIG_OCC_CONICAL SRCWA_EIG_OCC
#define SRCWA_EIG_OCC_CROSSED SRCWA_EIG_OCC
#endif

// sweeps the eigen kernels' deferred-update queue holds, for orders up to 512 and above
#ifndef SRCWA_EIG_QUEUE
#define SRCWA_EIG_QUEUE 4
#endif
#ifndef SRCWA_EIG_QUEUE_LARGE
#define SRCWA_EIG_QUEUE_LARGE 4
#endif
// orders above 256: columns per lane of the wave that runs each stage of the left rotation chain (a stage
// is 64 SRCWA_EIG_STAGE_LC rotations), and the rows that wave prefetches ahead; 0 (default) publishes every
// rotation through LDS with a workgroup barrier instead.  Measured on the 512-pillar batch (orders 242 to
// 722): staged chain with 2 columns per lane 8.19 s of eigensolves against 7.36 s for the barrier form --
// the chain's dependent rotation latency, not the barrier, sets the time
#ifndef SRCWA_EIG_STAGE_LC
#define SRCWA_EIG_STAGE_LC 0
#endif
#ifndef SRCWA_EIG_STAGE_RING
#define SRCWA_EIG_STAGE_RING 4
#endif
// 1: the mixed-precision star products (FP32 LU + FP64 refinement, wg_la32.hpp) are compiled into the
// S-matrix kernels and the batched mixed solve.  0 (default): compiled out and SRCWA_PLAN_MIXED refused.
// Their first GPU run ended in a memory fault whose cause is not found: the same code runs every shape of
// that run cleanly in a CPU emulation under AddressSanitizer (FP32 path and FP64 fallback), and it is the
// only code here that makes a real device function call.  Off until the cause is shown.
#ifndef SRCWA_MIXED_PRECISION
#define SRCWA_MIXED_PRECISION 0
#endif
// 1: the eigensolver accumulates its Householder reflectors into Z in blocks of 16 (two MFMA products per
// block) and runs the Hessenberg reduction in panels of 16 reflectors (three MFMA products per panel on the
// trailing matrix, orders from 72) instead of one reflector at a time.  0 (default): one at a time.  The
// blocked forms agree with the one-at-a-time forms in a CPU emulation of the workgroup (eigenvalues to
// 1e-15, eigen-pair residuals 4e-16 to 8e-16, clean under AddressSanitizer, including the matrix below),
// but their first GPU run ended in a memory fault on a planar layer matrix of order 101 (the reference's
// kRidgesSilver50TE table case), and with them the eigen kernels spill 1.2k-2k VGPRs (2.2-2.5 KB of
// scratch per lane against 0-44 B without); off until the cause is shown.
#ifndef SRCWA_EIG_BLOCKED_Z
#define SRCWA_EIG_BLOCKED_Z 0
#endif
#ifndef SRCWA_EIG_BLOCKED_REDUCE
#define SRCWA_EIG_BLOCKED_REDUCE 0
#endif
// minimum waves per SIMD each family's S-matrix kernel is built for (two: a few registers spill around
// phase boundaries, but a second instance per CU hides the LU and substitution latency chains)
#ifndef SRCWA_SMAT_OCC_PLANAR
#define SRCWA_SMAT_OCC_PLANAR 2
#endif
#ifndef SRCWA_SMAT_OCC_CONICAL
#define SRCWA_SMAT_OCC_CONICAL 2
#endif
#ifndef SRCWA_SMAT_OCC_CROSSED
#define SRCWA_SMAT_OCC_CROSSED 2
#endif
// SRCWA_SMAT_OCC (one value for every family) is what the tuner's variant builds set
#ifdef SRCWA_SMAT_OCC
#undef SRCWA_SMAT_OCC_PLANAR
#undef SRCWA_SMAT_OCC_CONICAL
#undef SRCWA_SMAT_OCC_CROSSED
#define SRCWA_SMAT_OCC_PLANAR SRCWA_SMAT_OCC
#define SRCWA_SMAT_OCC_CONICAL SRCWA_SMAT_OCC
#define SRCWA_SMAT_OCC_CROSSED SRCWA_SMAT_OCC
#endif

// 1: wg_lu and wg_lu_solve are called, not inlined, at each of their call sites (one copy per kernel
// instead of one per call site: a fraction of the compile time and code size of the S-matrix kernels)
#ifndef SRCWA_LU_OUTLINE
#define SRCWA_LU_OUTLINE 0
#endif
// 1: the QR rotation divides by a correctly rounded sqrt; 0: hardware reciprocal square root refined by
// two Newton steps.  Measured on MI355X over 112 matrices (random orders 41-241 and planar layer
// matrices): mean eigen-pair residual 6.2e-16 with 1 against 7.7e-16 with 0 (the reference QR: 6.5e-16),
// for 5 % more eigen time; so 1
#ifndef SRCWA_QR_EXACT_ROTATION
#define SRCWA_QR_EXACT_ROTATION 1
#endif

// (a minimum of one wave per SIMD is the same as no bound)
#define EIG_KERNEL_BOUNDS_OF(occ) __launch_bounds__(NT, ((occ) > 0 ? (occ) : 1))
#define SMAT_KERNEL_BOUNDS_OF(occ) __launch_bounds__(NT, ((occ) > 0 ? (occ) : 1))
