// oracle/ref_capi.cpp -- test infrastructure, never part of the product.
//
// A thin extern "C" wrapper around the reference's header-only C++17 solver, compiled from the headers
// where they lie (include/stablercwa/ of the reference tree) into oracle/_ref/libsrcwa_ref.so by
// oracle/Makefile.  Only tests/, the golden-vector scripts under tests/golden/ and bench.py's CPU
// baseline leg load it; it is the checker, never the thing measured as the GPU path.
//
// Stacks cross this boundary in the flat double encoding shared with the product C-ABI
// (include/srcwa.h, "flat stack records"):
//   1-D : period, nI.re, nI.im, nII.re, nII.im, L, then per layer
//         d, kind(0 uniform | 1 grating),
//           uniform: n.re, n.im
//           grating: K, edges[K+1], (n.re, n.im)[K], fourier_method(0 analytic | 1 FFT), fft_samples
//   2-D : Lx, Ly, nI, nII.re, nII.im, L, then per layer
//         d, kind(0 uniform | 1 patterned),
//           uniform: n.re, n.im
//           patterned: Ky, y_edges[Ky+1], then per y segment: Kx, x_edges[Kx+1], (n.re, n.im)[Kx]
#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "stablercwa/rcwa1d.hpp"
#include "stablercwa/rcwa1d_fields.hpp"
#include "stablercwa/rcwa1d_conical.hpp"
#include "stablercwa/rcwa1d_conical_fields.hpp"
#include "stablercwa/rcwa2d.hpp"
#include "stablercwa/eig.hpp"
#include "ref_decode.hpp"

namespace rcwa = stablercwa::rcwa;
using stablercwa::cplx;

namespace {

using oracle_decode::decode1d;
using oracle_decode::decode2d;

thread_local std::string g_error;

void put(double* out, const std::vector<cplx>& v) {
    for (std::size_t i = 0; i < v.size(); ++i) {
        out[2 * i] = v[i].real();
        out[2 * i + 1] = v[i].imag();
    }
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return -1;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -2;
    }
}

}  // namespace

namespace {
// Options of the *_ex solves: max_iter != 30 routes every layer eigensolve through eig_qr with that
// iteration limit (the reference tests' own idiom for an unconverged eigensolve, tests/test_rcwa1d.cpp:1169-1181,
// tests/test_rcwa1d_conical.cpp:554-563); diag = 1 sets diagnostics.  counts = {eig_info, n_eig, n_dense_solve,
// n_schur}; metrics = {max_eig_residual, max_solve_backward_error}.
template <class Opt>
void set_ex_options(Opt& opt, int max_iter, int diag) {
    opt.diagnostics = diag != 0;
    if (max_iter != 30)
        opt.custom_eig = [max_iter](const stablercwa::Matrix& a) {
            stablercwa::EigOptions e;
            e.max_iter = max_iter;
            return stablercwa::eig_qr(a, e);
        };
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

void ref_set_threads(int n) { stablercwa::set_num_threads(n); }

// One planar solve. Outputs have N = 2M+1 entries (R, T interleaved complex).
int ref_solve1d(const double* stack, double wavelength, double theta_deg, int M, int pol, int fast_path,
                double* R, double* T, double* DEr, double* DEt, int* eig_info, double* max_eig_residual) {
    return guarded([&] {
        rcwa::Options opt;
        opt.homogeneous_fast_path = fast_path != 0;
        opt.diagnostics = max_eig_residual != nullptr;
        const rcwa::Result r = rcwa::solve(decode1d(stack), wavelength, theta_deg, M,
                                           pol ? rcwa::Polarization::TM : rcwa::Polarization::TE, opt);
        put(R, r.R);
        put(T, r.T);
        std::memcpy(DEr, r.DEr.data(), r.DEr.size() * sizeof(double));
        std::memcpy(DEt, r.DEt.data(), r.DEt.size() * sizeof(double));
        *eig_info = r.eig_info;
        if (max_eig_residual) *max_eig_residual = r.max_eig_residual;
    });
}

// A wavelength sweep on the reference's thread pool (set by ref_set_threads). DEr/DEt: points x N.
int ref_sweep1d(const double* stack, int n_points, const double* wavelengths, const double* thetas, int M, int pol,
                double* DEr, double* DEt) {
    return guarded([&] {
        std::vector<rcwa::Point> pts(static_cast<std::size_t>(n_points));
        for (int i = 0; i < n_points; ++i) pts[i] = rcwa::Point{wavelengths[i], thetas[i]};
        const auto res = rcwa::sweep(decode1d(stack), pts, M, pol ? rcwa::Polarization::TM : rcwa::Polarization::TE);
        const int n = 2 * M + 1;
        for (int i = 0; i < n_points; ++i) {
            std::memcpy(DEr + static_cast<std::size_t>(i) * n, res[i].DEr.data(), n * sizeof(double));
            std::memcpy(DEt + static_cast<std::size_t>(i) * n, res[i].DEt.data(), n * sizeof(double));
        }
    });
}

// Near fields of one planar solve: rcwa::solve_with_fields, then field_at and field_map (rcwa1d_fields.hpp:55-206).
// available = 1: harm receives the 6 harmonic components (N x nz complex each, order i and depth iz at
// i + iz N) and, with nx > 0, grid the 6 grid components (nx x nz complex each, x_k and depth iz at
// k + iz nx); available = 0: the context is unavailable and nothing is written.  DEr, DEt: the solve's
// efficiencies (N each).
int ref_planar_fields(const double* stack, double wavelength, double theta_deg, int M, int pol, int nz,
                      const double* zs, int nx, double y, int* available, double* harm, double* grid, double* DEr,
                      double* DEt) {
    return guarded([&] {
        const auto res = rcwa::solve_with_fields(decode1d(stack), wavelength, theta_deg, M,
                                                 pol ? rcwa::Polarization::TM : rcwa::Polarization::TE);
        const int n = 2 * M + 1;
        std::memcpy(DEr, res.result.DEr.data(), n * sizeof(double));
        std::memcpy(DEt, res.result.DEt.data(), n * sizeof(double));
        *available = res.fields.available() ? 1 : 0;
        if (!*available) return;
        const stablercwa::RealVector z(zs, zs + nz);
        const auto h = rcwa::field_at(res, z);
        for (int c = 0; c < 6; ++c)
            for (int k = 0; k < n * nz; ++k) {
                const cplx v = h.components[c].data()[k];
                harm[2 * (static_cast<std::size_t>(c) * n * nz + k)] = v.real();
                harm[2 * (static_cast<std::size_t>(c) * n * nz + k) + 1] = v.imag();
            }
        if (nx <= 0) return;
        const auto g = rcwa::field_map(res, z, nx, y);
        for (int c = 0; c < 6; ++c)
            for (int k = 0; k < nx * nz; ++k) {
                const cplx v = g.components[c].data()[k];
                grid[2 * (static_cast<std::size_t>(c) * nx * nz + k)] = v.real();
                grid[2 * (static_cast<std::size_t>(c) * nx * nz + k) + 1] = v.imag();
            }
    });
}

// Near fields of one conical solve for the Jones vector jones = [Es.re, Es.im, Ep.re, Ep.im]:
// rcwa::solve_conical_with_fields, then field_at and field_map (rcwa1d_conical_fields.hpp:56-262).  Layout and
// availability as ref_planar_fields; n_schur: the solve's Schur-basis layers.
int ref_conical_fields(const double* stack, double wavelength, double theta_deg, double phi_deg, int M,
                       const double* jones, int nz, const double* zs, int nx, double y, int* available,
                       int* n_schur, double* harm, double* grid) {
    return guarded([&] {
        const auto res = rcwa::solve_conical_with_fields(decode1d(stack), wavelength, theta_deg, phi_deg, M);
        const int n = 2 * M + 1;
        *available = res.fields.available() ? 1 : 0;
        *n_schur = res.result.n_schur;
        if (!*available) return;
        const stablercwa::rcwa::JonesVector jv{{cplx(jones[0], jones[1]), cplx(jones[2], jones[3])}};
        const stablercwa::RealVector z(zs, zs + nz);
        const auto h = rcwa::field_at(res, z, jv);
        for (int c = 0; c < 6; ++c)
            for (int k = 0; k < n * nz; ++k) {
                const cplx v = h.components[c].data()[k];
                harm[2 * (static_cast<std::size_t>(c) * n * nz + k)] = v.real();
                harm[2 * (static_cast<std::size_t>(c) * n * nz + k) + 1] = v.imag();
            }
        if (nx <= 0) return;
        const auto g = rcwa::field_map(res, z, nx, y, jv);
        for (int c = 0; c < 6; ++c)
            for (int k = 0; k < nx * nz; ++k) {
                const cplx v = g.components[c].data()[k];
                grid[2 * (static_cast<std::size_t>(c) * nx * nz + k)] = v.real();
                grid[2 * (static_cast<std::size_t>(c) * nx * nz + k) + 1] = v.imag();
            }
    });
}

// One conical solve. Jr, Jt: N x 4 complex (ss, sp, ps, pp); DEr, DEt: N x 4 real.
int ref_solve_conical(const double* stack, double wavelength, double theta_deg, double phi_deg, int M,
                      double* Jr, double* Jt, double* DEr, double* DEt, int* eig_info, int* n_schur) {
    return guarded([&] {
        const rcwa::ConicalResult r = rcwa::solve_conical(decode1d(stack), wavelength, theta_deg, phi_deg, M);
        const std::size_t n = r.Jr.size();
        for (std::size_t j = 0; j < n; ++j) {
            const cplx a[4] = {r.Jr[j].ss, r.Jr[j].sp, r.Jr[j].ps, r.Jr[j].pp};
            const cplx b[4] = {r.Jt[j].ss, r.Jt[j].sp, r.Jt[j].ps, r.Jt[j].pp};
            for (int c = 0; c < 4; ++c) {
                Jr[8 * j + 2 * c] = a[c].real();
                Jr[8 * j + 2 * c + 1] = a[c].imag();
                Jt[8 * j + 2 * c] = b[c].real();
                Jt[8 * j + 2 * c + 1] = b[c].imag();
            }
            const double er[4] = {r.DEr[j].ss, r.DEr[j].sp, r.DEr[j].ps, r.DEr[j].pp};
            const double et[4] = {r.DEt[j].ss, r.DEt[j].sp, r.DEt[j].ps, r.DEt[j].pp};
            std::memcpy(DEr + 4 * j, er, sizeof er);
            std::memcpy(DEt + 4 * j, et, sizeof et);
        }
        *eig_info = r.eig_info;
        *n_schur = r.n_schur;
    });
}

// One 2-D solve. jones = [Es.re, Es.im, Ep.re, Ep.im]. DEr/DEt: N real; Rs, Rp, Ts, Tp: N complex.
int ref_solve2d(const double* stack, double wavelength, double theta_deg, double phi_deg, int M1, int M2,
                const double* jones, double* DEr, double* DEt, double* Rs, double* Rp, double* Ts, double* Tp,
                int* status, int* eig_info) {
    return guarded([&] {
        const rcwa::Jones2D j{cplx{jones[0], jones[1]}, cplx{jones[2], jones[3]}};
        const rcwa::Result2D r = rcwa::solve2d(decode2d(stack), wavelength, theta_deg, phi_deg, M1, M2, j);
        std::memcpy(DEr, r.DEr.data(), r.DEr.size() * sizeof(double));
        std::memcpy(DEt, r.DEt.data(), r.DEt.size() * sizeof(double));
        put(Rs, r.Rs);
        put(Rp, r.Rp);
        put(Ts, r.Ts);
        put(Tp, r.Tp);
        *status = static_cast<int>(r.status);
        *eig_info = r.eig_info;
    });
}

int ref_solve1d_ex(const double* stack, double wavelength, double theta_deg, int M, int pol, int max_iter, int diag,
                   double* R, double* T, double* DEr, double* DEt, int* counts, double* metrics) {
    return guarded([&] {
        rcwa::Options opt;
        set_ex_options(opt, max_iter, diag);
        const rcwa::Result r = rcwa::solve(decode1d(stack), wavelength, theta_deg, M,
                                           pol ? rcwa::Polarization::TM : rcwa::Polarization::TE, opt);
        put(R, r.R);
        put(T, r.T);
        std::memcpy(DEr, r.DEr.data(), r.DEr.size() * sizeof(double));
        std::memcpy(DEt, r.DEt.data(), r.DEt.size() * sizeof(double));
        counts[0] = r.eig_info;
        counts[1] = r.n_eig;
        counts[2] = r.n_dense_solve;
        counts[3] = 0;
        metrics[0] = r.max_eig_residual;
        metrics[1] = r.max_solve_backward_error;
    });
}

int ref_solve_conical_ex(const double* stack, double wavelength, double theta_deg, double phi_deg, int M,
                         int max_iter, int diag, double* Jr, double* Jt, double* DEr, double* DEt, int* counts,
                         double* metrics) {
    return guarded([&] {
        rcwa::Options opt;
        set_ex_options(opt, max_iter, diag);
        const rcwa::ConicalResult r = rcwa::solve_conical(decode1d(stack), wavelength, theta_deg, phi_deg, M, opt);
        const std::size_t n = r.Jr.size();
        for (std::size_t j = 0; j < n; ++j) {
            const cplx a[4] = {r.Jr[j].ss, r.Jr[j].sp, r.Jr[j].ps, r.Jr[j].pp};
            const cplx b[4] = {r.Jt[j].ss, r.Jt[j].sp, r.Jt[j].ps, r.Jt[j].pp};
            for (int c = 0; c < 4; ++c) {
                Jr[8 * j + 2 * c] = a[c].real();
                Jr[8 * j + 2 * c + 1] = a[c].imag();
                Jt[8 * j + 2 * c] = b[c].real();
                Jt[8 * j + 2 * c + 1] = b[c].imag();
            }
            const double er[4] = {r.DEr[j].ss, r.DEr[j].sp, r.DEr[j].ps, r.DEr[j].pp};
            const double et[4] = {r.DEt[j].ss, r.DEt[j].sp, r.DEt[j].ps, r.DEt[j].pp};
            std::memcpy(DEr + 4 * j, er, sizeof er);
            std::memcpy(DEt + 4 * j, et, sizeof et);
        }
        counts[0] = r.eig_info;
        counts[1] = r.n_eig;
        counts[2] = r.n_dense_solve;
        counts[3] = r.n_schur;
        metrics[0] = r.max_eig_residual;
        metrics[1] = r.max_solve_backward_error;
    });
}

int ref_solve2d_ex(const double* stack, double wavelength, double theta_deg, double phi_deg, int M1, int M2,
                   const double* jones, int max_iter, int diag, double* DEr, double* DEt, double* Rs, double* Rp,
                   double* Ts, double* Tp, int* status, int* counts, double* metrics) {
    return guarded([&] {
        rcwa::Options2D opt;
        set_ex_options(opt, max_iter, diag);
        const rcwa::Jones2D j{cplx{jones[0], jones[1]}, cplx{jones[2], jones[3]}};
        const rcwa::Result2D r = rcwa::solve2d(decode2d(stack), wavelength, theta_deg, phi_deg, M1, M2, j, opt);
        std::memcpy(DEr, r.DEr.data(), r.DEr.size() * sizeof(double));
        std::memcpy(DEt, r.DEt.data(), r.DEt.size() * sizeof(double));
        put(Rs, r.Rs);
        put(Rp, r.Rp);
        put(Ts, r.Ts);
        put(Tp, r.Tp);
        *status = static_cast<int>(r.status);
        counts[0] = r.eig_info;
        counts[1] = r.n_eig;
        counts[2] = r.n_dense_solve;
        counts[3] = 0;
        metrics[0] = r.max_eig_residual;
        metrics[1] = 0.0;
    });
}

// Many independent conical solves, each with its own stack, spread over the reference's thread pool with
// its own parallel_for (as sweep_conical spreads points, rcwa1d_conical.hpp:992-999).  Per instance:
// energy[i] = the sum of every channel efficiency (R + T of the s-input plus the p-input), eig_info[i].
// Used to time the reference on a batch of the benchmark's uneven instances.
int ref_solve_conical_many(int n, const double* const* stacks, const double* wavelengths, const double* thetas,
                           const double* phis, const int* Ms, double* energy, int* eig_info) {
    return guarded([&] {
        stablercwa::parallel_for(0, n, [&](int i) {
            const rcwa::ConicalResult r = rcwa::solve_conical(decode1d(stacks[i]), wavelengths[i], thetas[i], phis[i],
                                                              Ms[i]);
            double e = 0.0;
            for (std::size_t j = 0; j < r.DEr.size(); ++j)
                e += r.DEr[j].ss + r.DEr[j].sp + r.DEr[j].ps + r.DEr[j].pp + r.DEt[j].ss + r.DEt[j].sp + r.DEt[j].ps +
                     r.DEt[j].pp;
            energy[i] = e;
            eig_info[i] = r.eig_info;
        });
    });
}

// The same for 2-D solves: energy[i] = sum(DEr) + sum(DEt), status[i] = Status2D.  jones: 4 doubles per instance.
int ref_solve2d_many(int n, const double* const* stacks, const double* wavelengths, const double* thetas,
                     const double* phis, const int* M1s, const int* M2s, const double* jones, double* energy,
                     int* status) {
    return guarded([&] {
        stablercwa::parallel_for(0, n, [&](int i) {
            const double* jj = jones + 4 * i;
            const rcwa::Jones2D j{cplx{jj[0], jj[1]}, cplx{jj[2], jj[3]}};
            const rcwa::Result2D r = rcwa::solve2d(decode2d(stacks[i]), wavelengths[i], thetas[i], phis[i], M1s[i],
                                                   M2s[i], j);
            double e = 0.0;
            for (double v : r.DEr) e += v;
            for (double v : r.DEt) e += v;
            energy[i] = e;
            status[i] = static_cast<int>(r.status);
        });
    });
}

// eig_qr on a column-major n x n complex matrix. w: n complex; V: n x n complex; residual: eig_residual.
int ref_eig(int n, const double* A, double* w, double* V, int* info, double* residual) {
    return guarded([&] {
        stablercwa::Matrix a(n, n);
        std::memcpy(a.data(), A, sizeof(double) * 2 * n * n);
        const stablercwa::EigResult r = stablercwa::eig_qr(a);
        put(w, r.values);
        if (r.info == 0) std::memcpy(V, r.vectors.data(), sizeof(double) * 2 * n * n);
        *info = r.info;
        if (residual) *residual = r.info == 0 ? stablercwa::eig_residual(a, r) : -1.0;
    });
}

// The layer matrix the reference hands to its eigensolver (TE: Kx^2 - [[eps]], TM: the inverse-rule form).
int ref_layer_matrix(const double* stack, int layer, double wavelength, double theta_deg, int M, int pol, double* A) {
    return guarded([&] {
        const rcwa::Stack s = decode1d(stack);
        const stablercwa::Matrix a = rcwa::layer_matrix(s.layers.at(static_cast<std::size_t>(layer)), s.period,
                                                        wavelength, s.nI, theta_deg, M,
                                                        pol ? rcwa::Polarization::TM : rcwa::Polarization::TE);
        std::memcpy(A, a.data(), sizeof(double) * 2 * a.rows() * a.cols());
    });
}

}  // extern "C"
