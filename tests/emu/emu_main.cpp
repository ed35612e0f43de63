// Host run of the workgroup device routines of stablercwa_amd/csrc under the CPU emulation of
// tests/emu/hip/hip_runtime.h (one workgroup = 256 host threads).  Built with AddressSanitizer by
// tests/test_emulated_workgroup.py, so an out-of-bounds access in the emulated code aborts the run.
//
//   emu_main eig QC n seed [special]   eigendecompose a seeded random n x n matrix with wg_eig<QC>
//                                      (special: two diagonal blocks, so the reduction meets a zero column
//                                      below the sub-diagonal and then a zero leading entry); prints
//                                      info, the largest eigen-pair residual and the sorted eigenvalues
//   emu_main eigfile QC n path         the same for a column-major complex matrix read from a file
//   emu_main refine n nrhs seed [nan]  the mixed-precision solve (wg_refine_solve) of a seeded diagonally
//                                      weighted system; prints fallback, corrections and the backward error
//                                      (nan: its scratch starts as NaNs, as uncleared device memory may)
#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>

#include "hip/hip_runtime.h"

thread_local EmuDim3 threadIdx;
EmuDim3 blockIdx;
namespace emu {
std::barrier<>* block_bar;
std::barrier<>* wave_bar[4];
double xbuf_d[4][64];
long long xbuf_l[4][64];
double mf_a[4][64], mf_b[4][64];
double mf_c[4][64][4];
}  // namespace emu

#include "eig.hpp"
#if SRCWA_MIXED_PRECISION
#include "wg_la32.hpp"
#endif

using C = std::complex<double>;

template <class F>
static void run_block(F body) {
    std::barrier<> bb(256), w0(64), w1(64), w2(64), w3(64);
    emu::block_bar = &bb;
    emu::wave_bar[0] = &w0;
    emu::wave_bar[1] = &w1;
    emu::wave_bar[2] = &w2;
    emu::wave_bar[3] = &w3;
    std::vector<std::thread> ts;
    for (int t = 0; t < 256; ++t)
        ts.emplace_back([t, &body] {
            threadIdx.x = t;
            body();
        });
    for (auto& t : ts) t.join();
}

static C at(const std::vector<cd>& m, int i, int j, int n) { return C(m[i + (long)j * n].x, m[i + (long)j * n].y); }

template <int QC>
static int eig(int n, std::vector<cd> A) {
    const std::vector<cd> A0 = A;
    std::vector<cd> w(n), Z((long)n * n), X((long)n * n), ort(n), tmp(n);
    int info = -1;
    run_block([&] {
        const int r = wg_eig<QC>(n, A.data(), w.data(), Z.data(), X.data(), EigScratch{ort.data(), tmp.data()}, 30, kEps);
        if (threadIdx.x == 0) info = r;
    });
    double anorm = 0.0, worst = 0.0;
    for (const cd& v : A0) anorm += v.x * v.x + v.y * v.y;
    anorm = std::sqrt(anorm);
    for (int j = 0; j < n && info == 0; ++j) {
        double num = 0.0, den = 0.0;
        for (int i = 0; i < n; ++i) {
            C s = 0.0;
            for (int k = 0; k < n; ++k) s += at(A0, i, k, n) * at(A, k, j, n);
            s -= C(w[j].x, w[j].y) * at(A, i, j, n);
            num += std::norm(s);
            den += std::norm(at(A, i, j, n));
        }
        worst = std::max(worst, std::sqrt(num) / (anorm * std::sqrt(den)));
    }
    std::vector<C> ws(n);
    for (int i = 0; i < n; ++i) ws[i] = C(w[i].x, w[i].y);
    std::sort(ws.begin(), ws.end(), [](C a, C b) { return a.real() != b.real() ? a.real() < b.real() : a.imag() < b.imag(); });
    std::printf("info %d\nresidual %.17g\n", info, worst);
    for (const C& z : ws) std::printf("w %.17g %.17g\n", z.real(), z.imag());
    return 0;
}

static int eig_dispatch(int qc, int n, const std::vector<cd>& A) {
    switch (qc) {
    case 1: return eig<1>(n, A);
    case 2: return eig<2>(n, A);
    case 4: return eig<4>(n, A);
    default: std::fprintf(stderr, "QC %d not built into the emulator\n", qc); return 2;
    }
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "eig" && argc >= 5) {
        const int qc = std::atoi(argv[2]), n = std::atoi(argv[3]);
        std::mt19937_64 rng(std::strtoull(argv[4], nullptr, 10));
        std::normal_distribution<double> nd;
        std::vector<cd> A((long)n * n);
        for (cd& v : A) v = cd{nd(rng), nd(rng)};
        if (argc >= 6) {
            const int h = n / 2;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    if ((i < h) != (j < h)) A[i + (long)j * n] = cd{0.0, 0.0};
            A[(h + 1) + (long)h * n] = cd{0.0, 0.0};
        }
        return eig_dispatch(qc, n, A);
    }
    if (mode == "eigfile" && argc >= 5) {
        const int qc = std::atoi(argv[2]), n = std::atoi(argv[3]);
        std::vector<cd> A((long)n * n);
        FILE* f = std::fopen(argv[4], "rb");
        if (!f || std::fread(A.data(), sizeof(cd), A.size(), f) != A.size()) return 3;
        std::fclose(f);
        return eig_dispatch(qc, n, A);
    }
#if SRCWA_MIXED_PRECISION
    if (mode == "refine" && argc >= 5) {
        const int n = std::atoi(argv[2]), nrhs = std::atoi(argv[3]);
        std::mt19937_64 rng(std::strtoull(argv[4], nullptr, 10));
        std::normal_distribution<double> nd;
        std::vector<cd> A((long)n * n), B((long)n * nrhs), S(((n + 15) & ~15) + (((long)n * n + 15) & ~15L) +
                                                                refine_words(n, nrhs));
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) A[i + (long)j * n] = cd{nd(rng) + (i == j ? std::sqrt(double(n)) : 0.0), nd(rng)};
        for (cd& v : B) v = cd{nd(rng), nd(rng)};
        // "nan": the scratch starts as quiet NaNs instead of zeros (device memory is not cleared), so a
        // read of scratch that was never written shows in the results
        if (argc >= 6 && std::string(argv[5]) == "nan")
            for (cd& v : S) v = cd{std::nan(""), std::nan("")};
        const std::vector<cd> B0 = B;
        cd* lu64 = S.data();
        int* piv = reinterpret_cast<int*>(lu64 + (long)n * n);
        cd* scr = lu64 + (long)n * n + (n + 15) / 16 * 16;
        RefineStats st{};
        run_block([&] {
            const RefineStats s = wg_refine_solve(n, A.data(), B.data(), nrhs, scr, lu64, piv);
            if (threadIdx.x == 0) st = s;
        });
        double r2 = 0.0, a2 = 0.0, x2 = 0.0, b2 = 0.0;
        for (int j = 0; j < nrhs; ++j)
            for (int i = 0; i < n; ++i) {
                C s = -at(B0, i, j, n);
                for (int k = 0; k < n; ++k) s += at(A, i, k, n) * at(B, k, j, n);
                r2 += std::norm(s);
                x2 += std::norm(at(B, i, j, n));
                b2 += std::norm(at(B0, i, j, n));
            }
        for (const cd& v : A) a2 += v.x * v.x + v.y * v.y;
        std::printf("singular %d\nfallback %d\ncorrections %d\nberr %.17g\n", st.singular, st.fallback, st.iters,
                    std::sqrt(r2) / (std::sqrt(a2) * std::sqrt(x2) + std::sqrt(b2)));
        return 0;
    }
#endif
    return 2;
}
