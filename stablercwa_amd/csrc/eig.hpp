// Batched complex nonsymmetric eigensolver for gfx950: one 256-thread workgroup per matrix.
//
// The algorithm is the reference's eig_qr (include/stablercwa/eig_qr.hpp:1008-1124), i.e. EISPACK
// CORTH + COMQR2 without balancing:
//   1. Householder reduction to upper Hessenberg form            (eig_qr.hpp:231-272)
//   2. accumulate the reflectors into Z                          (eig_qr.hpp:354-378)
//   3. clear below the sub-diagonal, make the sub-diagonal real  (eig_qr.hpp:420-439)
//   4. single-shift QR sweeps with complex-cosine / real-sine rotations, Wilkinson shift with
//      power-of-two scaling, exceptional shifts at iterations 10 and 20, deflation
//      |h(l,l-1)| <= eps (|h(l-1,l-1)| + |h(l,l)|), max_iter sweeps per eigenvalue
//                                                                (eig_qr.hpp:442-607, 1061-1095)
//   5. back-substitution on T with the overflow control and the denominator floor
//      rel * max(|l_i|, |l_j|), V = Z X, columns normalised to unit 2-norm  (eig_qr.hpp:866-982)
//
// How the work maps onto the workgroup:
//   * Householder steps: one wave per column for the left reflector (lane-strided dot products),
//     one thread per row for the right reflector.
//   * QR sweep, left rotations: one thread per column.  The rotation chain G_{l+1..en} is serial;
//     the thread owning column i-1 forms G_i from the carried h(i-1,i-1) and the real h(i,i-1),
//     publishes it in LDS, and after one barrier every thread applies it to its own columns.
//     Each column streams down its rows with the running value in a register.
//   * QR sweep, right rotations on H and Z: one thread per row; a row passes through all the
//     sweep's rotations in order with the running value in a register, so H and Z rows are
//     independent and the column-major loads are coalesced across threads.
//   * Back-substitution: one thread per eigenvector, X kept row-major so the per-step updates are
//     coalesced; V = Z X on MFMA (wg_gemm with X^T), then a wave per column normalises.
#pragma once

#include "batch.hpp"
#include "tuned.hpp"
#include "wg_la.hpp"

constexpr int EIG_MAX_N = EIG_MAX_N_HOST;   // largest order the LDS rotation table holds
constexpr int EIG_RING = SRCWA_EIG_RING;    // rows each column prefetches ahead in the QR sweep

// Columns per lane of one wave for matrices of order n (QC): the eigen kernels are instantiated
// for 1, 2, 4, 8, 12 and 16 so small matrices do not pay the registers and LDS of the largest (12: orders
// up to 768, whose rotation table leaves room for two workgroups per CU; the crossed gratings' 722).
inline int eig_columns_per_thread(int n) {
    return n <= 64 ? 1 : n <= 128 ? 2 : n <= 256 ? 4 : n <= 512 ? 8 : n <= 768 ? 12 : 16;
}

// Translation-unit split.  The QC-templated kernels of a family dominate its compile time, so each
// family source is compiled once without SRCWA_QC_ONLY (its other kernels and its host launcher, the
// "main" unit) and once per QC in {1, 2, 4, 8, 12, 16} with SRCWA_QC_ONLY = QC (that QC's template kernels
// only); the units build in parallel.  The main unit reaches the per-QC units through plain host
// functions NAME_q1 .. NAME_q16 that take one argument struct, declared by SRCWA_QC_DISPATCH together
// with the dispatcher NAME(qc, args).
// The heaviest non-template kernels (the S-matrix recursions) likewise sit in a unit of their own,
// "part 1" (-DSRCWA_PART_ONLY=1), reached through a host function.
#define SRCWA_CAT2(a, b) a##b
#define SRCWA_CAT(a, b) SRCWA_CAT2(a, b)
#if defined(SRCWA_QC_ONLY)
#define SRCWA_UNIT_PART (-1)
#define SRCWA_QC_FN(NAME) SRCWA_CAT(NAME##_q, SRCWA_QC_ONLY)
#elif defined(SRCWA_PART_ONLY)
#define SRCWA_UNIT_PART SRCWA_PART_ONLY
#else
#define SRCWA_UNIT_PART 0
#endif
#define SRCWA_MAIN_TU (SRCWA_UNIT_PART == 0)
#define SRCWA_IN_PART(k) (SRCWA_UNIT_PART == (k))
#define SRCWA_QC_DISPATCH(NAME, ARGS)                                                     \
    void NAME##_q1(const ARGS&);                                                          \
    void NAME##_q2(const ARGS&);                                                          \
    void NAME##_q4(const ARGS&);                                                          \
    void NAME##_q8(const ARGS&);                                                          \
    void NAME##_q12(const ARGS&);                                                         \
    void NAME##_q16(const ARGS&);                                                         \
    inline void NAME(int qc, const ARGS& a) {                                             \
        switch (qc) {                                                                     \
            case 1: NAME##_q1(a); break;                                                  \
            case 2: NAME##_q2(a); break;                                                  \
            case 4: NAME##_q4(a); break;                                                  \
            case 8: NAME##_q8(a); break;                                                  \
            case 12: NAME##_q12(a); break;                                                \
            default: NAME##_q16(a); break;                                                \
        }                                                                                 \
    }

// Arguments of a family's QC-templated launches (which kernel, grid, stream and the kernel arguments;
// each family uses the fields its kernels take).
struct QcLaunch {
    int what;                 // family-specific kernel selector
    int grid;
    hipStream_t stream;
    int n_items;
    const void* insts;        // Inst1D* or Inst2D*
    const struct LayerTask* tasks;
    const double* stacks;
    const int64_t* layouts;
    cd* ws;
    int* task_info;
    int max_iter;
    const struct DenseItem* items;
    int* info;
    cd* wbase;                // dense eigensolves: base of the eigenvalue offsets (DenseItem::w)
    cd* sbase;                // dense eigensolves: base of the scratch offsets (DenseItem::scratch)
    const int* order;         // layer eigensolves: the launch's tasks, longest first (block b solves task order[b])
    const struct LayerOps* op_descs;   // conical modes: the plan's shared operator sets and their buffer
    cd* ops;
};

struct EigScratch {
    cd* ort;     // n
    cd* tmp;     // n
};

// Rotation G_i from the carried diagonal (xr + i xi) and the real sub-diagonal sub: c = x / |(x, sub)|,
// s = sub / |(x, sub)| (complex cosine, real sine), and the new diagonal |(x, sub)|.  Operands far from 1
// are scaled by a power of two first.  By default (SRCWA_QR_EXACT_ROTATION) the norm is a correctly
// rounded sqrt and c, s are correctly rounded quotients, as the reference's rotation (eig_qr.hpp:460-463);
// the alternative, the hardware reciprocal square root refined by two Newton steps, is 5 % faster but its
// rotations are less exactly unitary and the eigen-pair residuals come out 20 % higher.  Both sweep forms
// use this one function, so a matrix gets the same rotations whichever instantiation decomposes it.
DEV double qr_rotation(double xr, double xi, double sub, double& cr, double& ci, double& sn) {
    const double mm = fmax(fmax(fabs(xr), fabs(xi)), fabs(sub));
    int e = 0;
    if (!(mm > 1e-140 && mm < 1e140)) {
        e = binary_exponent(mm);
        xr = ldexp(xr, -e);
        xi = ldexp(xi, -e);
        sub = ldexp(sub, -e);
    }
    const double ss = fma(sub, sub, fma(xi, xi, xr * xr));
#if SRCWA_QR_EXACT_ROTATION
    const double nrm = sqrt(ss);
    cr = xr / nrm;
    ci = xi / nrm;
    sn = sub / nrm;
#else
    double inv = __builtin_amdgcn_rsq(ss);
    inv = fma(0.5 * inv, fma(-ss * inv, inv, 1.0), inv);
    inv = fma(0.5 * inv, fma(-ss * inv, inv, 1.0), inv);
    const double nrm = ss * inv;
    cr = xr * inv;
    ci = xi * inv;
    sn = sub * inv;
#endif
    return e ? ldexp(nrm, e) : nrm;
}

// Left rotation of one column's rows (i-1, i): top = conj(c) y + s z, bottom = c z - s y.  Written with
// explicit fused multiply-adds so both sweep forms evaluate it identically.
DEV void rotate_left(double cr, double ci, double sn, cd y, cd z, cd& top, cd& bottom) {
    top = cd{fma(sn, z.x, fma(ci, y.y, cr * y.x)), fma(sn, z.y, fma(-ci, y.x, cr * y.y))};
    bottom = cd{fma(-sn, y.x, fma(-ci, z.y, cr * z.x)), fma(-sn, y.y, fma(ci, z.x, cr * z.y))};
}

// The left half of one QR sweep, run by a single wave: rotations G_{l+1..en} (complex cosine, real
// sine, each formed from the diagonal carry of column i-1 and the real sub-diagonal h(i,i-1)) applied to
// rows l..en of columns l..n-1, the rotations stored in rot_c / rot_ci / rot_s for the right half.
// Lane c owns columns l + c + 64 q (q < LC); each column's next RD rows are loaded ahead into registers
// (rotation i only touches rows i-1 and i, so later rows are final when prefetched).  The chain stays
// inside one wave: the owner lane's values are broadcast with readlane, with no LDS round trip and no
// workgroup barrier per rotation.  Requires n <= 64 LC; the caller orders memory afterwards.
template <int LC, int RD, bool ROWMAJOR = false>
DEV void wave_left_sweep(int n, int l, int en, cd* __restrict__ H, double* rot_c, double* rot_ci, double* rot_s,
                         int cend = -1) {
    // columns l .. cend-1 are rotated (all of them by default; the workgroup solver defers the rest)
    if (cend < 0) cend = n;
    const int lane = lane_id();
    auto h = [&](int i, int j) -> cd& { return ROWMAJOR ? H[i * n + j] : H[i + j * n]; };
    cd carry[LC];
    cd ring[LC][RD];
#pragma unroll
    for (int q = 0; q < LC; ++q) {
        const int c = l + lane + 64 * q;
        if (c < cend) {
            carry[q] = h(l, c);
#pragma unroll
            for (int u = 0; u < RD; ++u)
                if (l + 1 + u <= en) ring[q][u] = h(l + 1 + u, c);
        }
    }
    for (int ib = l + 1; ib <= en; ib += RD) {
#pragma unroll
        for (int u = 0; u < RD; ++u) {
            const int i = ib + u;
            if (i > en) break;
            cd cur[LC];
#pragma unroll
            for (int q = 0; q < LC; ++q) {
                const int c = l + lane + 64 * q;
                cur[q] = ring[q][u];
                if (i + RD <= en && c >= i + RD - 1 && c < cend) ring[q][u] = h(i + RD, c);
            }
            // G_i from the owner column i-1: (h(i-1,i-1), h(i,i-1)) broadcast from its lane
            const int owner = i - 1 - l;
            const int oq = owner >> 6, ol = owner & 63;
            cd xc{0.0, 0.0}, zc{0.0, 0.0};
#pragma unroll
            for (int q = 0; q < LC; ++q)
                if (q == oq) {
                    xc = carry[q];
                    zc = cur[q];
                }
            double cr, ci, sn;
            const double nrm = qr_rotation(read_lane(xc.x, ol), read_lane(xc.y, ol), read_lane(zc.x, ol), cr, ci, sn);
            if (lane == ol) {
                h(i - 1, i - 1) = cd{nrm, 0.0};
                h(i, i - 1) = cd{0.0, 0.0};
            }
            if (lane == 0) {
                rot_c[i] = cr;
                rot_ci[i] = ci;
                rot_s[i] = sn;
            }
#pragma unroll
            for (int q = 0; q < LC; ++q) {
                const int c = l + lane + 64 * q;
                if (c >= i && c < cend) {
                    cd top;
                    rotate_left(cr, ci, sn, carry[q], cur[q], top, carry[q]);
                    h(i - 1, c) = top;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < LC; ++q) {
        const int c = l + lane + 64 * q;
        if (c >= en && c < cend) h(en, c) = carry[q];
    }
}

// The left half of one QR sweep with every thread of the workgroup: thread t owns columns
// l + t + NT q (q < CPT); the owner of column i-1 forms G_i, publishes it in LDS, and after a barrier
// every thread applies it to its own columns.  Each column's next RD rows are loaded ahead into
// registers.  H is row-major.  For large orders, where one wave would hold too many columns.
template <int CPT, int RD>
DEV void block_left_sweep(int n, int l, int en, cd* __restrict__ H, double* rot_c, double* rot_ci, double* rot_s,
                          int cend) {
    const int tid = threadIdx.x;
    auto h = [&](int i, int j) -> cd& { return H[i * n + j]; };
    cd carry[CPT];
    cd ring[CPT][RD];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int c = l + tid + q * NT;
        if (c < cend) {
            carry[q] = h(l, c);
#pragma unroll
            for (int u = 0; u < RD; ++u)
                if (l + 1 + u <= en) ring[q][u] = h(l + 1 + u, c);
        }
    }
    for (int ib = l + 1; ib <= en; ib += RD) {
#pragma unroll
        for (int u = 0; u < RD; ++u) {
            const int i = ib + u;
            if (i > en) break;
            cd cur[CPT];
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int c = l + tid + q * NT;
                cur[q] = ring[q][u];
                if (i + RD <= en && c >= i + RD - 1 && c < cend) ring[q][u] = h(i + RD, c);
            }
            // owner of column i-1 forms rotation G_i from (h(i-1,i-1), h(i,i-1))
            const int owner = i - 1 - l;
            if ((owner % NT) == tid) {
                const int qo = owner / NT;
                cd xc{0.0, 0.0}, zc{0.0, 0.0};
#pragma unroll
                for (int q = 0; q < CPT; ++q)
                    if (q == qo) {
                        xc = carry[q];
                        zc = cur[q];
                    }
                double cr, ci, sn;
                const double nrm = qr_rotation(xc.x, xc.y, zc.x, cr, ci, sn);
                h(i - 1, i - 1) = cd{nrm, 0.0};
                h(i, i - 1) = cd{0.0, 0.0};
                rot_c[i] = cr;
                rot_ci[i] = ci;
                rot_s[i] = sn;
            }
            __syncthreads();
            const double cr = rot_c[i], ci = rot_ci[i], sn = rot_s[i];
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const int c = l + tid + q * NT;
                if (c >= i && c < cend) {
                    cd top;
                    rotate_left(cr, ci, sn, carry[q], cur[q], top, carry[q]);
                    h(i - 1, c) = top;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int c = l + tid + q * NT;
        if (c >= en && c < cend) h(en, c) = carry[q];
    }
    __syncthreads();
}

// The left half of one QR sweep for large orders, in stages of B = 64 LC rotations.  A stage's rotations
// G_{i0..i1} pivot on columns i0-1 .. i1-1, so wave 0 runs the stage's chain on the B columns starting
// at i0-1 exactly as wave_left_sweep does (owner lane broadcast by readlane, no barrier per rotation),
// leaving each column's running value in row i1; then every thread applies the stage's B rotations,
// read from LDS, to one column right of that block at a time (rows i0-1 .. i1, loads batched eight
// ahead).  Two barriers per stage instead of one per rotation; each element sees the same rotations in
// the same order with the same arithmetic as in the other sweep forms.  H is row-major.
template <int LC, int RD>
DEV void staged_left_sweep(int n, int l, int en, cd* __restrict__ H, double* rot_c, double* rot_ci, double* rot_s,
                           int cend) {
    constexpr int B = 64 * LC;
    constexpr int LB = 8;
    for (int i0 = l + 1; i0 <= en; i0 += B) {
        const int i1 = min(i0 + B - 1, en);
        const int c0 = i0 - 1;
        const int cnear = min(c0 + B, cend);
        if (wave_id() == 0) wave_left_sweep<LC, RD, true>(n, c0, i1, H, rot_c, rot_ci, rot_s, cnear);
        __syncthreads();
        for (int c = cnear + threadIdx.x; c < cend; c += NT) {
            cd* col = H + c;
            cd carry = col[(long)c0 * n];
            for (int ib = i0; ib <= i1; ib += LB) {
                cd cur[LB];
#pragma unroll
                for (int u = 0; u < LB; ++u)
                    if (ib + u <= i1) cur[u] = ldc(col + (long)(ib + u) * n);
#pragma unroll
                for (int u = 0; u < LB; ++u) {
                    const int i = ib + u;
                    if (i <= i1) {
                        cd top;
                        rotate_left(rot_c[i], rot_ci[i], rot_s[i], carry, cur[u], top, carry);
                        stc(col + (long)(i - 1) * n, top);
                    }
                }
            }
            stc(col + (long)i1 * n, carry);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------
// Deferred updates of the QR sweeps.
//
// A sweep on the active window [l, en] only has to update H(l..en, l..en) before the next sweep can
// start (its shift, deflation test and rotations read nothing else).  The rest of its work -- the
// right rotations on Z and on the rows of H above l, the left rotations on the columns of H right of
// en, and the phase of column / row en there -- is queued (rotations and windows in LDS) and applied
// every few sweeps in one streaming pass: each row of Z, row of H above a window or column of H right
// of a window passes its elements through all queued sweeps in order, as a wavefront of one running
// value per sweep, so every element is loaded and stored once per pass instead of once per sweep.
// The sweeps are applied in their original order with the original arithmetic, so the result is the
// same as applying each sweep in full at once.  A pass is due when the queue is full, when the next
// window reaches above the top of a queued one (only then could the window touch queued work), and
// before the eigenvalue loop ends.

// Sweeps the queue holds (tuned): four cut the deferred traffic fourfold; more stages need more
// registers than the occupancy-tuned eigen kernels have (each stage holds two running values).
__host__ __device__ constexpr int eig_queue_sweeps(int QC) { return QC <= 8 ? SRCWA_EIG_QUEUE : SRCWA_EIG_QUEUE_LARGE; }

struct EigQueue {
    double* rot;       // [3 KQ][stride]: per queued sweep s, rows 3s, 3s+1, 3s+2 = c, ci, s of rotation j
    int* lo;           // window of each queued sweep
    int* hi;
    double* phase;     // [2 KQ] phase of column en (1 when none)
    int stride;        // n + 1
};

// One stream (a row for right rotations, a column for left rotations): elements k0..k1 at base[k * step]
// pass through the queued sweeps whose bit is set in act, in order, as a systolic pipeline.  Stage s
// holds the running value of sweep s; at step k it takes the element its predecessor emitted at step
// k - 1 (element k - 2s; stage 0 takes the loaded element k) and emits the element before it.  All
// stages of one step are independent, so their rotations overlap.  Outside its window a stage rotates
// by the identity (c = 1, s = 0: emit the held value, hold the new one), i.e. it only delays.
template <int KQ, bool LEFT, bool ALL = false>
DEV void eig_stream(cd* base, long step, int nq, unsigned act, const EigQueue& q) {
    // the windows are the same for every lane (scalar registers); act says which apply to this stream
    int lo[KQ], hi[KQ];
    int k0 = 1 << 30, k1 = -1;
#pragma unroll
    for (int s = 0; s < KQ; ++s) {
        lo[s] = __builtin_amdgcn_readfirstlane(s < nq ? q.lo[s] : (1 << 30));
        hi[s] = __builtin_amdgcn_readfirstlane(s < nq ? q.hi[s] : -2);
        if (s >= nq) act &= ~(1u << s);
        if ((act >> s) & 1u) {
            k0 = min(k0, lo[s]);
            k1 = max(k1, hi[s]);
        }
    }
    if (k1 < 0) return;
    constexpr int D = 4;                // elements loaded ahead (a shift register of loads in flight)
    constexpr int LAG = 2 * KQ - 1;     // the last stage emits element k - LAG at step k
    cd pend[KQ], carry[KQ], ahead[D];
#pragma unroll
    for (int s = 0; s < KQ; ++s) {
        pend[s] = cd{0.0, 0.0};
        carry[s] = cd{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < D; ++u) ahead[u] = k0 + u <= k1 ? ldc(base + (k0 + u) * step) : cd{0.0, 0.0};
    for (int k = k0; k <= k1 + LAG; ++k) {
        keep_loads_in_loop();     // one step's rotation coefficients at a time
        const cd x = ahead[0];
#pragma unroll
        for (int u = 0; u + 1 < D; ++u) ahead[u] = ahead[u + 1];
        ahead[D - 1] = k + D <= k1 ? ldc(base + (k + D) * step) : cd{0.0, 0.0};
        // last stage first, so each stage still reads its predecessor's output of the previous step
#pragma unroll
        for (int s = KQ - 1; s >= 0; --s) {
            const int j = k - 2 * s;
            const bool on = s < nq && (ALL || ((act >> s) & 1u));
            // a sweep's table holds the identity outside its window (and every index is clamped into
            // the table), so only a stage that is unused or does not apply to this stream needs a select
            const int jc = min(max(j, 0), q.stride - 1);
            const double* r = q.rot + 3 * s * q.stride + jc;
            double cr = r[0], ci = r[q.stride], sn = r[2 * q.stride];
            if (!on) {
                cr = 1.0;
                ci = 0.0;
                sn = 0.0;
            }
            const cd a = pend[s], b = s == 0 ? x : carry[s - 1];
            cd out;
            if (LEFT) {
                out = cd{cr * a.x + ci * a.y + sn * b.x, cr * a.y - ci * a.x + sn * b.y};
                pend[s] = cd{cr * b.x - ci * b.y - sn * a.x, cr * b.y + ci * b.x - sn * a.y};
            } else {
                out = cd{cr * a.x - ci * a.y + sn * b.x, cr * a.y + ci * a.x + sn * b.y};
                pend[s] = cd{cr * b.x + ci * b.y - sn * a.x, cr * b.y - ci * b.x - sn * a.y};
            }
            if (on && j - 1 == hi[s]) {
                // the end of sweep s's chain: its phase (none is stored as exactly 1)
                const cd ph{q.phase[2 * s], q.phase[2 * s + 1]};
                if (ph.y != 0.0) out = LEFT ? conj(ph) * out : out * ph;
            }
            carry[s] = out;
        }
        const int jo = k - LAG;
        if (jo >= k0 && jo <= k1) stc(base + jo * step, carry[KQ - 1]);
    }
}

// Apply the nq queued sweeps: every row of Z, the rows of H above the queued windows and the columns of
// H right of them.  H is row-major (X), Z column-major.  Ends with a barrier.
template <int KQ>
DEV void eig_flush(int n, cd* __restrict__ HR, cd* __restrict__ Z, int nq, const EigQueue& q) {
    __syncthreads();
    int lmax = 0, emin = n;
    for (int s = 0; s < nq; ++s) {
        lmax = max(lmax, q.lo[s]);
        emin = min(emin, q.hi[s]);
    }
    const int nrow = lmax, ncol = n - 1 - emin;
    for (int t = threadIdx.x; t < n + nrow + ncol; t += NT) {
        unsigned act = 0;
        if (t < n + nrow) {
            // right rotations: a row of Z (column-major) or a row of H above the windows (row-major)
            const bool zrow = t < n;
            const int r = zrow ? t : t - n;
            for (int s = 0; s < nq; ++s)
                if (zrow || r < q.lo[s]) act |= 1u << s;
            if (zrow) eig_stream<KQ, false, true>(Z + r, n, nq, act, q);
            else eig_stream<KQ, false>(HR + (long)r * n, 1, nq, act, q);
        } else {
            // left rotations: a column of H right of the windows
            const int c = emin + 1 + (t - n - nrow);
            for (int s = 0; s < nq; ++s)
                if (c > q.hi[s]) act |= 1u << s;
            eig_stream<KQ, true>(HR + c, n, nq, act, q);
        }
    }
    __syncthreads();
}

// Householder reflector applied to columns j0..n-1 of M (rows r0 .. r0+len-1, column-major, leading
// dimension ld): d_j = sum_k conj(u_k) M(r0+k, j), then M(r0+k, j) += t(d_j) u_k.  One thread per column,
// so no reduction crosses lanes; each pass loads eight entries at a time (the column is back in the
// caches for the second pass).  u is in LDS.
template <class TFun>
DEV void reflect_columns(cd* M, long ld, int r0, int len, int j0, int n, const cd* u, TFun tfun) {
    constexpr int B = 8;
    for (int j = j0 + threadIdx.x; j < n; j += NT) {
        cd* a = M + r0 + (long)j * ld;
        // B partial sums (entry k goes to sum k mod B), added pairwise at the end: about as accurate as
        // a tree reduction over the column, and B independent chains
        double pr[B], pi[B];
#pragma unroll
        for (int b = 0; b < B; ++b) pr[b] = pi[b] = 0.0;
        int k = 0;
        for (; k + B <= len; k += B) {
            cd av[B];
#pragma unroll
            for (int b = 0; b < B; ++b) av[b] = ldc(a + k + b);
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const cd uk = u[k + b];
                pr[b] += uk.x * av[b].x + uk.y * av[b].y;
                pi[b] += uk.x * av[b].y - uk.y * av[b].x;
            }
        }
        for (int b = 0; k < len; ++k, ++b) {
            const cd uk = u[k], ak = ldc(a + k);
            // the tail (fewer than B entries) continues the same B sums
#pragma unroll
            for (int q = 0; q < B; ++q)
                if (q == b) {
                    pr[q] += uk.x * ak.x + uk.y * ak.y;
                    pi[q] += uk.x * ak.y - uk.y * ak.x;
                }
        }
#pragma unroll
        for (int w = B / 2; w > 0; w >>= 1)
#pragma unroll
            for (int b = 0; b < w; ++b) {
                pr[b] += pr[b + w];
                pi[b] += pi[b + w];
            }
        const cd t = tfun(pr[0], pi[0]);
        k = 0;
        for (; k + B <= len; k += B) {
            cd av[B];
#pragma unroll
            for (int b = 0; b < B; ++b) av[b] = ldc(a + k + b);
#pragma unroll
            for (int b = 0; b < B; ++b) stc(a + k + b, av[b] + t * u[k + b]);
        }
        for (; k < len; ++k) stc(a + k, ldc(a + k) + t * u[k]);
    }
    __syncthreads();
}

// Reflector blocks of the blocked Householder steps.
constexpr int HB = 16;
// Orders from which the blocked forms run (their scratch, 3 n HB words, must fit in one n x n matrix).
constexpr int HB_MIN_N = 3 * HB;

// Z = P_1 ... P_{n-2} (the CORTH reflectors, step 2) in blocks of HB reflectors, last block first: the block
// P_{i0} ... P_{i1} = I + V T V^H (compact WY form, T upper triangular from the recurrence
// T(0:k, k) = tau_k T(0:k, 0:k) V(:, 0:k)^H v_k, T(k, k) = tau_k) is applied to Z(i0:n, i0:n) as two MFMA
// products, W = V^H Z, Z += V (T W).  Reflector i is v = 2^-e (ort[i], h(i+1, i-1), ..., h(n-1, i-1)) with
// tau = 1 / Re(conj(2^-e h(i, i-1)) v_0), the same reflector as the one-at-a-time form builds (a skipped
// reflector is the identity, tau = 0).  scratch: 3 n HB words (V n x HB, V^H and W HB x n); t: HB x HB in LDS.
DEV void accumulate_reflectors_blocked(int n, const cd* __restrict__ H, const cd* __restrict__ ort, cd* __restrict__ Z,
                                       cd* __restrict__ scratch, cd* __restrict__ t) {
    const int tid = threadIdx.x;
    cd* V = scratch;                        // n x HB, column-major
    cd* VH = scratch + (long)n * HB;        // HB x n, V^H
    cd* W = scratch + 2 * (long)n * HB;     // HB x n
    __shared__ double s_tau[HB];
    __shared__ int s_exp[HB];
    for (int i1 = n - 2; i1 >= 1; i1 -= HB) {
        const int i0 = max(1, i1 - HB + 1), b = i1 - i0 + 1;
        const int rows = n - i0;
        // reflector scalings, one thread per reflector
        if (tid < b) {
            const int i = i0 + tid;
            const cd oi = ort[i], sub0 = H[i + (long)(i - 1) * n];
            double tau = 0.0;
            int e = 0;
            if (!is_zero(oi) && !is_zero(sub0)) {
                e = binary_exponent(abs1(oi));
                const cd sub = scale2(sub0, -e), v0 = scale2(oi, -e);
                tau = 1.0 / (sub.x * v0.x + sub.y * v0.y);
            }
            s_tau[tid] = tau;
            s_exp[tid] = e;
        }
        __syncthreads();
        // V (rows i0..n-1) and V^H
        for (int e = tid; e < rows * b; e += NT) {
            const int r = i0 + e % rows, k = e / rows, i = i0 + k;
            cd v{0.0, 0.0};
            if (s_tau[k] != 0.0 && r >= i) v = scale2(r == i ? ort[i] : H[r + (long)(i - 1) * n], -s_exp[k]);
            V[r + (long)k * n] = v;
            VH[k + (long)r * HB] = conj(v);
        }
        __syncthreads();
        // G = V^H V into W (HB x HB), then T by one thread
        wg_gemm<GEMM_SET, false>(b, b, rows, VH + (long)i0 * HB, HB, V + i0, n, W, HB);
        if (tid == 0) {
            for (int k = 0; k < b; ++k) {
                const double tau = s_tau[k];
                for (int r = 0; r < k; ++r) {
                    cd s{0.0, 0.0};
                    for (int l = r; l < k; ++l) s = s + t[r + l * HB] * W[l + k * HB];
                    t[r + k * HB] = tau * s;
                }
                t[k + k * HB] = cd{tau, 0.0};
                for (int r = k + 1; r < b; ++r) t[r + k * HB] = cd{0.0, 0.0};
            }
        }
        __syncthreads();
        // W = V^H Z(i0:n, i0:n);  W = T W (one column per thread);  Z(i0:n, i0:n) += V W
        wg_gemm<GEMM_SET, false>(b, rows, rows, VH + (long)i0 * HB, HB, Z + i0 + (long)i0 * n, n, W, HB);
        for (int c = tid; c < rows; c += NT) {
            cd w[HB];
#pragma unroll
            for (int k = 0; k < HB; ++k) w[k] = k < b ? W[k + (long)c * HB] : cd{0.0, 0.0};
#pragma unroll
            for (int r = 0; r < HB; ++r) {
                cd s{0.0, 0.0};
#pragma unroll
                for (int k = 0; k < HB; ++k)
                    if (k >= r && k < b) s = s + t[r + k * HB] * w[k];
                if (r < b) W[r + (long)c * HB] = s;
            }
        }
        __syncthreads();
        wg_gemm<GEMM_ADD, false>(rows, rows, b, V + i0, n, W, HB, Z + i0 + (long)i0 * n, n);
    }
}

// Orders from which the blocked reduction runs (its scratch, 4 n HB + 2 HB words, fits in one n x n matrix).
constexpr int HB_MIN_N_REDUCE = 72;

// Step 1 (CORTH) in panels of HB reflectors.  Within a panel, step m = m0 + k needs column c = m - 1 of
// Q_k^H A Q_k (Q_k = P_{m0} ... P_{m-1} = I - V T V^H, compact WY form; A the matrix at the panel's start):
// its right part is A(:, c) - Y conj(V(c, :))^T with Y = A V T, its left part then subtracts
// V T^H V^H (rows >= m0).  From that column the step forms the reflector exactly as the one-at-a-time
// reduction (scale, u, hh, the modified first entry, the new sub-diagonal, ort[m], and the column's entries
// below the sub-diagonal left as they were), writes the now final column c, and extends
// Y by y_k = tau_k (A(:, m:n) u_k - Y V^H u_k) and T by T(0:k, k) = -tau_k T V^H u_k.  After the panel the
// trailing columns J >= m0 + b - 1 get A Q (A -= Y V^H) and then Q^H A (A -= V T^H (V^H A)), three MFMA
// products.  Every column of the matrix is read once per reflector (for y_k) instead of the two read-write
// passes per side of the one-at-a-time form.  wcol: n words; scratch: 4 n HB + 2 HB words; t: HB x HB in LDS.
DEV void reduce_hessenberg_blocked(int n, cd* __restrict__ A, cd* __restrict__ ort, cd* __restrict__ wcol,
                                   cd* __restrict__ scratch, cd* __restrict__ t) {
    const int tid = threadIdx.x;
    cd* V = scratch;                        // n x HB
    cd* VH = V + (long)n * HB;              // HB x n
    cd* Y = VH + (long)n * HB;              // n x HB
    cd* W = Y + (long)n * HB;               // HB x n
    cd* sv = W + (long)n * HB;              // HB
    cd* pv = sv + HB;                       // HB
    __shared__ cd s_sp[HB];
    auto a = [&](int i, int j) -> cd& { return A[i + (long)j * n]; };
    for (int m0 = 1; m0 <= n - 2; m0 += HB) {
        const int b = min(HB, n - 1 - m0);
        for (int e = tid; e < n * HB; e += NT) {
            V[e] = cd{0.0, 0.0};
            VH[e] = cd{0.0, 0.0};
        }
        __syncthreads();
        for (int k = 0; k < b; ++k) {
            const int m = m0 + k, c = m - 1;
            // (a) the column with the panel's earlier right reflectors applied
            for (int i = tid; i < n; i += NT) {
                cd v = a(i, c);
                for (int j = 0; j < k; ++j) v = v - Y[i + (long)j * n] * VH[j + (long)c * HB];
                wcol[i] = v;
            }
            __syncthreads();
            // (b) and the earlier left reflectors (rows >= m0): wcol -= V T^H (V^H wcol)
            if (k > 0) {
                wg_gemm<GEMM_SET, false>(k, 1, n - m0, VH + (long)m0 * HB, HB, wcol + m0, n, sv, HB);
                if (tid < k) {
                    cd s{0.0, 0.0};
                    for (int j = 0; j <= tid; ++j) s = s + conj(t[j + tid * HB]) * sv[j];
                    s_sp[tid] = s;
                }
                __syncthreads();
                for (int i = m0 + tid; i < n; i += NT) {
                    cd v = wcol[i];
                    for (int j = 0; j < k; ++j) v = v - V[i + (long)j * n] * s_sp[j];
                    wcol[i] = v;
                }
                __syncthreads();
            }
            // (c) the reflector from wcol(m:n), as the one-at-a-time reduction forms it
            double part = 0.0;
            for (int i = m + tid; i < n; i += NT) part += abs1(wcol[i]);
            const double scale = block_sum(part);
            double tau = 0.0;
            if (scale == 0.0) {
                if (tid == 0) ort[m] = cd{0.0, 0.0};
                for (int i = tid; i < n; i += NT) a(i, c) = wcol[i];
            } else {
                double hp = 0.0;
                for (int i = m + tid; i < n; i += NT) {
                    const cd u = cdiv(wcol[i], scale);
                    V[i + (long)k * n] = u;
                    hp += norm2(u);
                }
                double hh = block_sum(hp);          // also the barrier: V(:, k) is complete
                double g = sqrt(hh);
                const cd u0 = V[m + (long)k * n];
                const double f = cabs(u0);
                cd sub, first;
                if (f == 0.0) {
                    first = cd{g, 0.0};
                    sub = cd{-g * scale, 0.0};
                } else {
                    hh += f * g;
                    g /= f;
                    first = u0 * (1.0 + g);
                    sub = wcol[m] * (-g);
                }
                tau = 1.0 / hh;
                __syncthreads();                     // every thread has read V(m, k)
                if (tid == 0) {
                    V[m + (long)k * n] = first;
                    ort[m] = first * scale;
                }
                // the final column c: rows above m as updated, the new sub-diagonal, the entries below left
                for (int i = tid; i < n; i += NT) a(i, c) = i == m ? sub : wcol[i];
                __syncthreads();
                for (int i = m + tid; i < n; i += NT) VH[k + (long)i * HB] = conj(V[i + (long)k * n]);
            }
            __syncthreads();
            // (d) p = V^H u_k,  y_k = tau (A(:, m:n) u_k - Y p)
            if (k > 0) wg_gemm<GEMM_SET, false>(k, 1, n - m, VH + (long)m * HB, HB, V + m + (long)k * n, n, pv, HB);
            for (int i = tid; i < n; i += NT) {
                cd y{0.0, 0.0};
                if (tau != 0.0) {
                    for (int j = m; j < n; ++j) y = y + a(i, j) * V[j + (long)k * n];
                    for (int l = 0; l < k; ++l) y = y - Y[i + (long)l * n] * pv[l];
                    y = tau * y;
                }
                Y[i + (long)k * n] = y;
            }
            // (e) T(0:k, k) = -tau T(0:k, 0:k) p,  T(k, k) = tau
            if (tid == 0) {
                for (int r = 0; r < k; ++r) {
                    cd s{0.0, 0.0};
                    for (int l = r; l < k; ++l) s = s + t[r + l * HB] * pv[l];
                    t[r + k * HB] = -tau * s;
                }
                t[k + k * HB] = cd{tau, 0.0};
                for (int r = k + 1; r < HB; ++r) t[r + k * HB] = cd{0.0, 0.0};
            }
            __syncthreads();
        }
        // trailing columns J0..n-1:  A -= Y V^H;  then rows m0..n-1:  A -= V T^H (V^H A)
        const int J0 = m0 + b - 1, nj = n - J0;
        wg_gemm<GEMM_SUB, false>(n, nj, b, Y, n, VH + (long)J0 * HB, HB, A + (long)J0 * n, n);
        wg_gemm<GEMM_SET, false>(b, nj, n - m0, VH + (long)m0 * HB, HB, A + m0 + (long)J0 * n, n, W, HB);
        for (int col = tid; col < nj; col += NT) {
            cd w[HB];
#pragma unroll
            for (int j = 0; j < HB; ++j) w[j] = j < b ? W[j + (long)col * HB] : cd{0.0, 0.0};
#pragma unroll
            for (int r = 0; r < HB; ++r) {
                cd s{0.0, 0.0};
#pragma unroll
                for (int j = 0; j < HB; ++j)
                    if (j <= r) s = s + conj(t[j + r * HB]) * w[j];
                if (r < b) W[r + (long)col * HB] = s;
            }
        }
        __syncthreads();
        wg_gemm<GEMM_SUB, false>(n - m0, nj, b, V + m0, n, W, HB, A + m0 + (long)J0 * n, n);
    }
}

// Eigen-decompose the n x n matrix A (column-major, leading dimension n) in place.
//   in : A = the matrix (overwritten)
//   out: A = right eigenvectors (unit 2-norm columns) when info == 0; w = eigenvalues
// Z and X are n x n scratch matrices.  Returns info (0, or k > 0 when eigenvalue k did not converge).
// With vectors == false the Schur form is returned instead (the reference's keep_transform without
// vectors): A = T (upper triangular, eigenvalues on the diagonal), Z = the unitary transform,
// Z^H A_in Z = T.
template <int QC>
DEV int wg_eig(int n, cd* __restrict__ A, cd* __restrict__ w, cd* __restrict__ Z, cd* __restrict__ X,
               EigScratch s, int max_iter, double eps, bool vectors = true) {
    if (n > 64 * QC) return n + 1;     // the launch chose too few columns per lane
    const int tid = threadIdx.x;
    const int lane = lane_id(), wave = wave_id();
    cd* H = A;
    cd* ort = s.ort;
    auto h = [&](int i, int j) -> cd& { return H[i + (long)j * n]; };
    // One LDS block holds the current Householder vector and the blocked reflectors' T factor in steps 1-2
    // and the queued sweeps' rotations in step 4 (never live at once), so the largest instantiations fit
    // two workgroups per CU.
    constexpr int KQ = eig_queue_sweeps(QC);
    constexpr int QROT_WORDS = 3 * KQ * (64 * QC + 1), U_WORDS = 2 * 64 * QC + 2 * HB * HB;
    __shared__ __align__(16) double s_lds[QROT_WORDS > U_WORDS ? QROT_WORDS : U_WORDS];
    cd* const s_u = reinterpret_cast<cd*>(s_lds);

    // ---- 1. Householder reduction to Hessenberg form (CORTH)
    if (SRCWA_EIG_BLOCKED_REDUCE && n >= HB_MIN_N_REDUCE)
        reduce_hessenberg_blocked(n, H, ort, s.tmp, X, reinterpret_cast<cd*>(s_lds) + 64 * QC);
    else
    for (int m = 1; m < n - 1; ++m) {
        cd* x = H + (long)(m - 1) * n;
        double part = 0.0;
        for (int i = m + tid; i < n; i += NT) part += abs1(x[i]);
        const double scale = block_sum(part);
        if (scale == 0.0) {
            if (tid == 0) ort[m] = cd{0.0, 0.0};
            __syncthreads();
            continue;
        }
        double hp = 0.0;
        for (int i = m + tid; i < n; i += NT) {
            const cd u = cdiv(x[i], scale);
            ort[i] = u;
            hp += norm2(u);
        }
        double hh = block_sum(hp);          // also a barrier: ort[] is complete
        double g = sqrt(hh);
        const double f = cabs(ort[m]);
        if (f == 0.0) {
            __syncthreads();
            if (tid == 0) {
                ort[m] = cd{g, 0.0};
                x[m] = cd{scale, 0.0};
            }
        } else {
            hh += f * g;
            g /= f;
            const cd om = ort[m];
            __syncthreads();
            if (tid == 0) ort[m] = om * (1.0 + g);
        }
        __syncthreads();
        // the reflector in LDS for both sides
        const int len = n - m;
        for (int i = tid; i < len; i += NT) s_u[i] = ort[m + i];
        __syncthreads();
        // left: columns m..n-1, one thread per column
        reflect_columns(H, n, m, len, m, n, s_u, [&](double sr, double si) { return cd{-sr / hh, -si / hh}; });
        // right: rows 0..n-1, one thread per row (entries of a row are n apart, coalesced across rows)
        for (int i = tid; i < n; i += NT) {
            constexpr int B = 8;
            cd acc{0.0, 0.0};
            int j = n - 1;
            for (; j - B + 1 >= m; j -= B) {
                cd hv[B];
#pragma unroll
                for (int b = 0; b < B; ++b) hv[b] = ldc(H + i + (long)(j - b) * n);
#pragma unroll
                for (int b = 0; b < B; ++b) acc = acc + s_u[j - b - m] * hv[b];
            }
            for (; j >= m; --j) acc = acc + s_u[j - m] * ldc(H + i + (long)j * n);
            acc = cdiv(acc, hh);
            j = m;
            for (; j + B <= n; j += B) {
                cd hv[B];
#pragma unroll
                for (int b = 0; b < B; ++b) hv[b] = ldc(H + i + (long)(j + b) * n);
#pragma unroll
                for (int b = 0; b < B; ++b) stc(H + i + (long)(j + b) * n, hv[b] - conj(s_u[j + b - m]) * acc);
            }
            for (; j < n; ++j) stc(H + i + (long)j * n, ldc(H + i + (long)j * n) - conj(s_u[j - m]) * acc);
        }
        __syncthreads();
        if (tid == 0) {
            ort[m] = ort[m] * scale;
            x[m] = x[m] * (-g);
        }
        __syncthreads();
    }

    // ---- 2. Z = P_1 ... P_{n-2}
    wg_identity(n, Z, n);
    if (SRCWA_EIG_BLOCKED_Z && n >= HB_MIN_N) {
        accumulate_reflectors_blocked(n, H, ort, Z, X, reinterpret_cast<cd*>(s_lds) + 64 * QC);
    } else
    for (int i = n - 2; i >= 1; --i) {
        const cd oi = ort[i];
        const cd sub0 = h(i, i - 1);
        if (is_zero(oi) || is_zero(sub0)) continue;
        const int e = binary_exponent(abs1(oi));
        const cd sub = scale2(sub0, -e);
        const cd v0 = scale2(oi, -e);
        const double minus_h = sub.x * v0.x + sub.y * v0.y;
        const int len = n - i;
        for (int k = tid; k < len; k += NT) s_u[k] = k == 0 ? v0 : scale2(h(i + k, i - 1), -e);
        __syncthreads();
        reflect_columns(Z, n, i, len, i, n, s_u, [&](double sr, double si) { return cd{sr / minus_h, si / minus_h}; });
    }

    // ---- 3. clear below the sub-diagonal; make the sub-diagonal real (D^H H D, Z D)
    for (int e = tid; e < n * n; e += NT) {
        const int i = e % n, j = e / n;
        if (i >= j + 2) h(i, j) = cd{0.0, 0.0};
    }
    __syncthreads();
    // y_i (unit phases) form a serial chain through the column scaling of step i-1
    cd* y = s.tmp;
    if (tid == 0) {
        cd prev{1.0, 0.0};
        bool prev_scaled = false;
        y[0] = cd{1.0, 0.0};
        for (int i = 1; i < n; ++i) {
            cd sub = h(i, i - 1);
            if (prev_scaled) sub = sub * prev;
            if (sub.y == 0.0) {
                h(i, i - 1) = sub;
                y[i] = cd{1.0, 0.0};
                prev_scaled = false;
            } else {
                double norm = 0.0;
                const cd yy = unit_phase(sub, norm);
                h(i, i - 1) = cd{norm, 0.0};
                y[i] = yy;
                prev = yy;
                prev_scaled = true;
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += NT) {
        const int i = e % n, j = e / n;
        if (i <= j) {
            cd val = h(i, j);
            if (i >= 1) val = conj(y[i]) * val;
            val = val * y[j];
            h(i, j) = val;
        }
        Z[e] = Z[e] * y[j];
    }
    __syncthreads();

    // ---- 4. shifted QR iterations (COMQR2)
    // H is iterated in row-major order in X (free until the back-substitution): the left rotations
    // walk rows with one thread per column, so each rotation's loads and stores are coalesced across
    // the threads; the right rotations of H walk each row from one thread.  Z stays column-major.
    for (int e = tid; e < n * n; e += NT) X[e] = H[(e / n) + (e % n) * n];
    __syncthreads();
    double* const s_qrot = s_lds;
    __shared__ int s_qlo[KQ], s_qhi[KQ];
    __shared__ double s_qphase[2 * KQ];
    __shared__ int s_info;
    const EigQueue queue{s_qrot, s_qlo, s_qhi, s_qphase, n + 1};
    int nq = 0, qtop = 0;          // queued sweeps; largest top row l among them
    cd t{0.0, 0.0};
    int en = n - 1, its = 0, info = 0;
    {
    cd* HR = X;
    auto h = [&](int i, int j) -> cd& { return HR[i * n + j]; };
    while (true) {
        int l = 0;
        bool sweep = false;
        if (en >= 0) {
            int lc = 0;
            for (int i = 1 + tid; i <= en; i += NT)
                if (abs1(h(i, i - 1)) <= eps * (abs1(h(i - 1, i - 1)) + abs1(h(i, i)))) lc = i;
            l = block_max_int(lc);
            if (l == en) {
                __syncthreads();
                if (tid == 0) {
                    const cd val = h(en, en) + t;
                    w[en] = val;
                    h(en, en) = val;
                }
                __syncthreads();
                --en;
                its = 0;
                continue;
            }
            if (its == max_iter) info = en + 1;
            else sweep = true;
        }
        // the queue is applied when full, when the next window reaches above a queued one (only then
        // could it touch pending work), and once the iteration ends (the only call site, so the pass
        // is compiled once)
        if (nq > 0 && (!sweep || nq == KQ || l < qtop)) {
            eig_flush<KQ>(n, HR, Z, nq, queue);
            nq = 0;
            qtop = 0;
        }
        if (!sweep) break;
        cd sh;
        if (its == 10 || its == 20) {
            sh = cd{fabs(h(en, en - 1).x) + (en >= 2 ? fabs(h(en - 1, en - 2).x) : 0.0), 0.0};
        } else {
            const cd a = h(en - 1, en - 1), b = h(en - 1, en), d = h(en, en);
            const double c = h(en, en - 1).x;
            const cd yv = 0.5 * (a - d);
            const int e = binary_exponent(fmax(fmax(abs1(b), fabs(c)), abs1(yv)));
            const cd xv = scale2(b, -e) * ldexp(c, -e);
            if (is_zero(xv)) {
                sh = d;
            } else {
                const cd ys = scale2(yv, -e);
                cd zz = csqrt(ys * ys + xv);
                if (ys.x * zz.x + ys.y * zz.y < 0.0) zz = -zz;
                sh = d - scale2(cdiv(xv, ys + zz), e);
            }
        }
        double* rot_c = s_qrot + 3 * nq * queue.stride;
        double* rot_ci = rot_c + queue.stride;
        double* rot_s = rot_ci + queue.stride;
        // the identity outside this sweep's window (the left sweep fills l+1..en)
        for (int j = tid; j <= n; j += NT)
            if (j <= l || j > en) {
                rot_c[j] = 1.0;
                rot_ci[j] = 0.0;
                rot_s[j] = 0.0;
            }
        __syncthreads();
        for (int i = tid; i <= en; i += NT) h(i, i) = h(i, i) - sh;
        t = t + sh;
        ++its;
        __syncthreads();

        // (1) left rotations.  Orders up to 256: the rotation chain runs inside wave 0 (QC columns per
        //     lane, rows of the row-major H loaded coalesced EIG_RING rows ahead) while the other waves
        //     wait at the barrier.  Larger orders: stages of 64 SRCWA_EIG_STAGE_LC rotations, each chain
        //     in wave 0 on the block of columns it pivots on, then the stage's rotations applied to the
        //     columns right of the block by every thread.
        //     Only columns l..en: the columns right of en are queued.
        if constexpr (QC <= 4) {
            if (wave == 0) wave_left_sweep<QC, EIG_RING, true>(n, l, en, HR, rot_c, rot_ci, rot_s, en + 1);
        } else {
#if SRCWA_EIG_STAGE_LC > 0
            staged_left_sweep<SRCWA_EIG_STAGE_LC, SRCWA_EIG_STAGE_RING>(n, l, en, HR, rot_c, rot_ci, rot_s, en + 1);
#else
            block_left_sweep<QC / 4, EIG_RING>(n, l, en, HR, rot_c, rot_ci, rot_s, en + 1);
#endif
        }
        __syncthreads();

        // (2) phase that makes r(en, en) real (row en right of en and column en above l are queued)
        const cd hee = h(en, en);
        const bool has_phase = hee.y != 0.0;
        cd phase{1.0, 0.0};
        if (has_phase) {
            double nrm = 0.0;
            phase = unit_phase(hee, nrm);
            __syncthreads();
            if (tid == 0) h(en, en) = cd{nrm, 0.0};
        }
        if (tid == 0) {
            s_qlo[nq] = l;
            s_qhi[nq] = en;
            s_qphase[2 * nq] = phase.x;
            s_qphase[2 * nq + 1] = phase.y;
        }
        ++nq;
        qtop = max(qtop, l);
        __syncthreads();

        // (3) right rotations on the rows l..en of H, one thread per row (rows above l and Z are
        //     queued).  Each row loads the next RB entries before rotating through them, so RB loads
        //     are in flight at once.
        constexpr int RB = 8;
        for (int row = l + tid; row <= en; row += NT) {
            cd* base = HR + (long)row * n;     // H row-major
            const int j0 = max(l + 1, row);
            cd a = base[j0 - 1];
            for (int jb = j0; jb <= en; jb += RB) {
                cd bv[RB];
#pragma unroll
                for (int u = 0; u < RB; ++u)
                    if (jb + u <= en) bv[u] = base[jb + u];
#pragma unroll
                for (int u = 0; u < RB; ++u) {
                    const int j = jb + u;
                    if (j <= en) {
                        const double cr = rot_c[j], ci = rot_ci[j], sn = rot_s[j];
                        const cd b = bv[u];
                        base[j - 1] = cd{cr * a.x - ci * a.y + sn * b.x, cr * a.y + ci * a.x + sn * b.y};
                        a = cd{cr * b.x + ci * b.y - sn * a.x, cr * b.y - ci * b.x - sn * a.y};
                    }
                }
            }
            if (has_phase) a = a * phase;
            base[en] = a;
        }
        __syncthreads();
    }
    }
    // back to column-major for the eigenvectors (or the Schur form the caller keeps)
    for (int e = tid; e < n * n; e += NT) H[(e / n) + (e % n) * n] = X[e];
    if (tid == 0) s_info = info;
    __syncthreads();
    info = s_info;
    if (info != 0) {
        for (int i = tid; i <= en; i += NT) w[i] = h(i, i) + t;
        __syncthreads();
        return info;
    }

    // ---- 5. eigenvectors of T, V = Z X, normalise
    for (int j = 0; j + 1 < n; ++j)
        if (tid == 0) h(j + 1, j) = cd{0.0, 0.0};
    __syncthreads();
    if (!vectors) return 0;
    double nm = 0.0;
    for (int e = tid; e < n * n; e += NT) {
        const int i = e % n, j = e / n;
        if (i <= j) nm = fmax(nm, abs1(H[e]));
    }
    const double tnorm = block_max(nm);
    const double rel = fmax(eps, kEps);
    double tiny = rel * tnorm;
    if (tiny == 0.0) tiny = 2.2250738585072014e-308;
    constexpr double kQuotientMax = 1e300;
    const double kShrink = ldexp(1.0, -512);
    // X row-major: X[k*n + j] = component k of eigenvector j
    for (int e = tid; e < n * n; e += NT) {
        const int k = e / n, j = e % n;
        X[e] = k < j ? h(k, j) : (k == j ? cd{1.0, 0.0} : cd{0.0, 0.0});
    }
    __syncthreads();
    for (int j = tid; j < n; j += NT) {
        const cd wj = w[j];
        for (int i = j - 1; i >= 0; --i) {
            const cd wi = w[i];
            cd yv = wj - wi;
            if (is_zero(yv)) {
                yv = cd{tiny, 0.0};
            } else {
                const double fl = rel * fmax(abs1(wj), abs1(wi));
                if (abs1(yv) < fl) yv = cd{fl, 0.0};
            }
            for (int pass = 0; pass < 3 && abs1(X[(long)i * n + j]) > abs1(yv) * kQuotientMax; ++pass)
                for (int k = 0; k <= j; ++k) X[(long)k * n + j] = X[(long)k * n + j] * kShrink;
            const cd xi = cdiv(X[(long)i * n + j], yv);
            X[(long)i * n + j] = xi;
            const double tr = abs1(xi);
            cd xs = xi;
            if (tr != 0.0 && tr + 1.0 / tr <= tr) {
                for (int k = 0; k <= j; ++k) X[(long)k * n + j] = cdiv(X[(long)k * n + j], tr);
                xs = X[(long)i * n + j];
            }
            const cd* ti = H + (long)i * n;
            for (int k = 0; k < i; ++k) X[(long)k * n + j] = X[(long)k * n + j] + xs * ti[k];
        }
    }
    __syncthreads();
    // V = Z X into A (T is no longer needed)
    wg_gemm<GEMM_SET, true>(n, n, n, Z, n, X, n, A, n);
    for (int j = wave; j < n; j += NWAVE) {
        cd* vj = A + (long)j * n;
        double big = 0.0;
        for (int i = lane; i < n; i += 64) big = fmax(big, fmax(fabs(vj[i].x), fabs(vj[i].y)));
        big = wave_max(big);
        if (big == 0.0 || !isfinite(big)) continue;
        double ssq = 0.0;
        for (int i = lane; i < n; i += 64) ssq += norm2(cdiv(vj[i], big));
        ssq = wave_sum(ssq);
        const double inv = 1.0 / sqrt(ssq);
        for (int i = lane; i < n; i += 64) vj[i] = cdiv(vj[i], big) * inv;
    }
    __syncthreads();
    return 0;
}
