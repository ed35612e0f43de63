// C++17 host API of the MI355X-native RCWA solver: the reference's rcwa::solve, sweep, solve_conical,
// sweep_conical and solve2d entry points with the same argument meaning, result fields and error
// behaviour, computed on the GPU through the C-ABI in srcwa.h (link with -lsrcwa_amd).
//
//   namespace rcwa = srcwa_amd::rcwa;
//   rcwa::Result r = rcwa::solve(stack, 1.053, 66.36, 20, rcwa::Polarization::TE);
//
// Reference signatures kept: rcwa1d.hpp:1077 (solve), :1084 (sweep); rcwa1d_conical.hpp:985 (solve_conical),
// :992 (sweep_conical); rcwa2d.hpp:1355 (solve2d).  Types mirror rcwa1d.hpp:80-164,
// rcwa1d_conical.hpp:81-183 and rcwa2d_types.hpp:27-130.
//
// Error behaviour: invalid input throws std::invalid_argument with the reference's message; a numerical
// failure returns NaN arrays with eig_info set (planar, conical) or status NumericalFailure (crossed);
// options this build does not provide (custom hooks -- see stablercwa_amd/hooks.hpp for the other
// direction --, LR eigensolver, non-pivoting or QR linear solvers, homogeneous_fast_path = false,
// crossed-grating near-field capture) throw srcwa_amd::Unsupported; a HIP runtime error throws
// std::runtime_error.  Planar and conical near fields are in stablercwa_amd/fields.hpp.  FFT-sampled
// grating layers (FourierMethod::FFT) are computed on the GPU.  sweep* run every point in one batched GPU
// launch; results keep the input order and each result is bitwise the same as its single solve.
#pragma once

#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <utility>
#include <string>
#include <vector>

#include "../srcwa.h"

namespace srcwa_amd {

using cplx = std::complex<double>;
using Vector = std::vector<cplx>;
using RealVector = std::vector<double>;

// An option or layer kind the reference accepts and this GPU build does not compute.
struct Unsupported : std::logic_error {
    using std::logic_error::logic_error;
};

namespace detail {

inline void check(int code) {
    if (code == SRCWA_OK) return;
    const std::string msg = srcwa_last_error();
    if (code == SRCWA_EINVAL) throw std::invalid_argument(msg);
    if (code == SRCWA_EUNSUPPORTED) throw Unsupported(msg);
    if (code == SRCWA_EUNAVAILABLE) throw std::logic_error(msg);
    throw std::runtime_error(msg);
}

// One context (HIP stream) per device and plan-flag set for the whole process (the flags are fixed per
// context, so concurrent callers with different options never share one); contexts live until the
// process ends.
inline srcwa_context* context(int device, int plan_flags = 0) {
    static std::mutex lock;
    static std::map<std::pair<int, int>, srcwa_context*> contexts;
    std::lock_guard<std::mutex> guard(lock);
    const auto key = std::make_pair(device, plan_flags);
    auto it = contexts.find(key);
    if (it != contexts.end()) return it->second;
    srcwa_context* ctx = nullptr;
    check(srcwa_context_create(device, &ctx));
    if (plan_flags) check(srcwa_context_set_plan_flags(ctx, plan_flags));
    contexts[key] = ctx;
    return ctx;
}

// Per-solve diagnostics of a plan's last execute (srcwa_plan_diagnostics) and its stage times.
struct PlanReport {
    std::vector<double> diag;     // 5 per problem
    std::vector<double> refine;   // 3 per problem (srcwa_plan_refinement)
    double ms[4] = {0, 0, 0, 0};  // setup + assembly, eigensolves, modes + S-matrix, total
};
inline PlanReport report(srcwa_plan* p, int problems) {
    PlanReport r;
    r.diag.resize(5 * static_cast<std::size_t>(problems));
    r.refine.resize(3 * static_cast<std::size_t>(problems));
    check(srcwa_plan_diagnostics(p, r.diag.data()));
    check(srcwa_plan_refinement(p, r.refine.data()));
    check(srcwa_plan_stage_ms(p, r.ms));
    return r;
}

// Plan flags of a call's options.
template <class Opt>
int plan_flags(const Opt& o) {
    return (o.diagnostics ? SRCWA_PLAN_DIAGNOSTICS : 0) | (o.mixed_precision ? SRCWA_PLAN_MIXED : 0) |
           (o.share_layer_operators ? 0 : SRCWA_PLAN_POINT_OPS);
}

// RAII owner of a launch plan.
struct Plan {
    srcwa_plan* p = nullptr;
    Plan() = default;
    Plan(const Plan&) = delete;
    Plan& operator=(const Plan&) = delete;
    ~Plan() {
        if (p) srcwa_plan_destroy(p);
    }
    void run() {
        check(srcwa_plan_execute(p));
        check(srcwa_plan_sync(p));
    }
};

inline std::vector<cplx> as_complex(const std::vector<double>& v, std::size_t first, std::size_t count) {
    std::vector<cplx> out(count);
    for (std::size_t i = 0; i < count; ++i) out[i] = cplx{v[2 * (first + i)], v[2 * (first + i) + 1]};
    return out;
}

constexpr double kPi = 3.14159265358979323846;

using Clock = std::chrono::steady_clock;
inline double seconds_since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

}  // namespace detail

namespace rcwa {

enum class Polarization { TE, TM };
enum class FourierMethod { Analytic, FFT };
enum class EigMethod { QR, LR };
enum class LinearSolver { LUPartial, LUNoPivot, QRHouseholder };

// Dense matrix handed to the reference's hooks (column-major complex); kept so option structs match.
struct Matrix {
    int rows = 0, cols = 0;
    std::vector<cplx> data;
};
struct EigResult {
    Vector values;
    Matrix vectors;
    int sweeps = 0;
    int info = 0;
};

struct Layer {
    double d = 0.0;
    cplx n = 1.0;
    std::vector<double> edges;
    std::vector<cplx> n_values;
    bool homogeneous = true;
    FourierMethod fourier_method = FourierMethod::Analytic;
    int fft_samples = 0;

    static Layer uniform(double thickness, cplx index) {
        Layer layer;
        layer.d = thickness;
        layer.n = index;
        return layer;
    }
    static Layer grating(double thickness, std::vector<double> boundaries, std::vector<cplx> indices,
                         FourierMethod method = FourierMethod::Analytic, int samples = 0) {
        Layer layer;
        layer.d = thickness;
        layer.edges = std::move(boundaries);
        layer.n_values = std::move(indices);
        layer.homogeneous = false;
        layer.fourier_method = method;
        layer.fft_samples = samples;
        return layer;
    }
};

struct Stack {
    cplx nI = 1.0;
    std::vector<Layer> layers;
    cplx nII = 1.0;
    double period = 1.0;
};

struct Options {
    EigMethod eig = EigMethod::QR;
    LinearSolver linear = LinearSolver::LUPartial;
    bool homogeneous_fast_path = true;
    bool diagnostics = false;          // eigen-pair residuals and solve backward errors, computed on the GPU
    bool parallel_layers = false;      // accepted; layers always run concurrently on the GPU
    std::function<EigResult(const Matrix&)> custom_eig;
    std::function<Matrix(const Matrix&, const Matrix&)> custom_multiply;
    std::function<Matrix(const Matrix&, const Matrix&)> custom_solve;
    // this build only: the GPU, the per-launch device workspace bound (0 = 64 GiB), the mixed-precision
    // star products (S-matrix solves by FP32 LU + FP64 iterative refinement; SRCWA_PLAN_MIXED), and whether
    // the points of a sweep share each layer's wavelength-invariant operators (false: every point computes
    // its own, SRCWA_PLAN_POINT_OPS; the results are the same bits)
    int device = 0;
    std::int64_t max_workspace_bytes = 0;
    bool mixed_precision = false;
    bool share_layer_operators = true;
};

struct Point {
    double wavelength = 1.0;
    double theta_deg = 0.0;
};

struct Timing {
    double total = 0, eig = 0, linear = 0, product = 0;
};

struct Result {
    std::vector<int> orders;
    Vector R, T;
    RealVector DEr, DEt;
    double reflectance() const {
        double sum = 0.0;
        for (double v : DEr) sum += v;
        return sum;
    }
    double transmittance() const {
        double sum = 0.0;
        for (double v : DEt) sum += v;
        return sum;
    }
    double max_eig_residual = 0;
    double max_solve_backward_error = 0;
    int eig_info = 0;
    int n_eig = 0, n_dense_solve = 0;
    Timing timing;
    // this build only (Options::mixed_precision): solves that fell back to FP64, FP32 correction steps,
    // largest final refined residual ||b - A x|| / (||A|| ||x||)
    int refine_fallbacks = 0, refine_iterations = 0;
    double refine_residual = 0;
};

// ---- conical ---------------------------------------------------------------------------------
struct Jones {
    cplx ss, sp, ps, pp;
};
struct JonesEfficiency {
    double ss = 0, sp = 0, ps = 0, pp = 0;
};

struct ConicalPolarized {
    cplx Es, Ep;
    Vector Rs, Rp, Ts, Tp;
    RealVector DEr, DEt;
    RealVector DEr_s, DEr_p, DEt_s, DEt_p;
    double reflectance() const {
        double sum = 0.0;
        for (double v : DEr) sum += v;
        return sum;
    }
    double transmittance() const {
        double sum = 0.0;
        for (double v : DEt) sum += v;
        return sum;
    }
};

struct ConicalResult {
    std::vector<int> orders;
    RealVector kx;
    double ky = 0;
    RealVector phi_m_deg;
    Vector kzI, kzII;
    RealVector flux_r, flux_t;
    std::vector<Jones> Jr, Jt;
    std::vector<JonesEfficiency> DEr, DEt;

    // Amplitudes and efficiencies for the input Jones vector [Es, Ep], normalised to unit length
    // (rcwa1d_conical.hpp:116-150).
    ConicalPolarized for_input(cplx Es, cplx Ep) const {
        const double len = std::hypot(std::abs(Es), std::abs(Ep));
        if (!(len > 0.0) || !std::isfinite(len))
            throw std::invalid_argument("rcwa: the input Jones vector must be finite and nonzero");
        ConicalPolarized p;
        p.Es = Es / len;
        p.Ep = Ep / len;
        const std::size_t n = Jr.size();
        for (auto* v : {&p.Rs, &p.Rp, &p.Ts, &p.Tp}) v->assign(n, cplx{});
        for (auto* v : {&p.DEr, &p.DEt, &p.DEr_s, &p.DEr_p, &p.DEt_s, &p.DEt_p}) v->assign(n, 0.0);
        for (std::size_t j = 0; j < n; ++j) {
            p.Rs[j] = Jr[j].ss * p.Es + Jr[j].sp * p.Ep;
            p.Rp[j] = Jr[j].ps * p.Es + Jr[j].pp * p.Ep;
            p.Ts[j] = Jt[j].ss * p.Es + Jt[j].sp * p.Ep;
            p.Tp[j] = Jt[j].ps * p.Es + Jt[j].pp * p.Ep;
            p.DEr_s[j] = flux_r[j] * std::norm(p.Rs[j]);
            p.DEr_p[j] = flux_r[j] * std::norm(p.Rp[j]);
            p.DEt_s[j] = flux_t[j] * std::norm(p.Ts[j]);
            p.DEt_p[j] = flux_t[j] * std::norm(p.Tp[j]);
            p.DEr[j] = p.DEr_s[j] + p.DEr_p[j];
            p.DEt[j] = p.DEt_s[j] + p.DEt_p[j];
        }
        return p;
    }
    // [Es, Ep] = [sin psi exp(i delta), cos psi] (rcwa1d_conical.hpp:151-155)
    ConicalPolarized for_angles(double psi_deg, double delta_deg = 0.0) const {
        const double psi = psi_deg * detail::kPi / 180.0, delta = delta_deg * detail::kPi / 180.0;
        return for_input(std::sin(psi) * cplx{std::cos(delta), std::sin(delta)}, cplx{std::cos(psi), 0.0});
    }
    // sum over orders of flux J^H J for both sides as [ss, sp, ps, pp]; the identity for lossless stacks
    std::array<cplx, 4> flux_unitarity() const {
        std::array<cplx, 4> u{};
        for (std::size_t j = 0; j < Jr.size(); ++j)
            for (int side = 0; side < 2; ++side) {
                const Jones& J = side == 0 ? Jr[j] : Jt[j];
                const double f = side == 0 ? flux_r[j] : flux_t[j];
                u[0] += f * (std::norm(J.ss) + std::norm(J.ps));
                u[1] += f * (std::conj(J.ss) * J.sp + std::conj(J.ps) * J.pp);
                u[2] += f * (std::conj(J.sp) * J.ss + std::conj(J.pp) * J.ps);
                u[3] += f * (std::norm(J.sp) + std::norm(J.pp));
            }
        return u;
    }

    double max_eig_residual = 0;
    double max_solve_backward_error = 0;
    int eig_info = 0;
    int n_eig = 0, n_dense_solve = 0;
    int n_schur = 0;
    Timing timing;
    int refine_fallbacks = 0, refine_iterations = 0;   // this build only, as Result
    double refine_residual = 0;
    double min_abs_lambda = std::numeric_limits<double>::infinity();
    bool degenerate_mode_warning = false;
};

struct ConicalPoint {
    double wavelength = 1.0;
    double theta_deg = 0.0;
    double phi_deg = 0.0;
};

// ---- crossed gratings ------------------------------------------------------------------------
enum class Factorization2D { Li, Laurent };
enum class Status2D { Ok, NumericalFailure, RayleighAnomaly };

struct XProfile2D {
    std::vector<double> edges;
    std::vector<cplx> n_values;
};

struct Layer2D {
    double d = 0.0;
    bool homogeneous = true;
    cplx n = 1.0;
    std::vector<double> y_edges;
    std::vector<XProfile2D> x_profiles;
    std::string name;

    static Layer2D uniform(double thickness, cplx index, std::string label = {}) {
        Layer2D layer;
        layer.d = thickness;
        layer.n = index;
        layer.name = std::move(label);
        return layer;
    }
    static Layer2D patterned(double thickness, std::vector<double> y_boundaries, std::vector<XProfile2D> rows,
                             std::string label = {}) {
        Layer2D layer;
        layer.d = thickness;
        layer.homogeneous = false;
        layer.y_edges = std::move(y_boundaries);
        layer.x_profiles = std::move(rows);
        layer.name = std::move(label);
        return layer;
    }
};

struct Stack2D {
    double nI = 1.0;
    cplx nII = 1.0;
    std::vector<Layer2D> layers;
    double Lx = 1.0, Ly = 1.0;
};

struct Options2D {
    EigMethod eig = EigMethod::QR;
    Factorization2D factorization = Factorization2D::Li;
    bool homogeneous_fast_path = true;
    bool diagnostics = false;
    bool parallel_layers = false;
    std::function<EigResult(const Matrix&)> custom_eig;
    std::function<Matrix(const Matrix&, const Matrix&)> custom_multiply;
    std::function<Matrix(const Matrix&, const Matrix&)> custom_solve;
    int device = 0;
    std::int64_t max_workspace_bytes = 0;
    bool mixed_precision = false;
    bool share_layer_operators = true;
};

struct Timing2D {
    double total = 0, operators = 0, eig = 0, linear = 0, product = 0;
};

using Jones2D = std::array<cplx, 2>;

inline Jones2D jones_from_angles_2d(double psi_deg = 90.0, double delta_deg = 0.0) {
    const double psi = psi_deg * detail::kPi / 180.0, delta = delta_deg * detail::kPi / 180.0;
    return {std::sin(psi) * cplx{std::cos(delta), std::sin(delta)}, cplx{std::cos(psi), 0.0}};
}

inline int harmonic_index_2d(int m, int n, int M1, int M2) { return (m + M1) * (2 * M2 + 1) + (n + M2); }

struct Result2D {
    Status2D status = Status2D::Ok;
    std::string message;
    int M1 = 0, M2 = 0, N = 0, i0 = 0;
    std::vector<int> m, n;
    double k0 = 0.0, theta = 0.0, phi = 0.0;
    RealVector Kx, Ky;
    Vector kzI, kzII;
    Vector Ex_r, Ey_r, Ez_r, Ex_t, Ey_t, Ez_t;
    Vector Rs, Rp, Ts, Tp;
    RealVector DEr, DEt;
    std::vector<char> prop_r, prop_t;
    double reflectance() const {
        double s = 0;
        for (double v : DEr) s += v;
        return s;
    }
    double transmittance() const {
        double s = 0;
        for (double v : DEt) s += v;
        return s;
    }
    int eig_info = 0;
    int n_eig = 0, n_dense_solve = 0;
    double max_eig_residual = 0;
    Timing2D timing;
    int refine_fallbacks = 0, refine_iterations = 0;   // this build only, as Result
    double refine_residual = 0;
};

namespace detail {

using namespace ::srcwa_amd::detail;

template <class Opt>
void require_gpu_options(const Opt& o) {
    if (o.eig != EigMethod::QR) throw Unsupported("rcwa: this build's eigensolver is the shifted QR (EigMethod::QR)");
    if (!o.homogeneous_fast_path) throw Unsupported("rcwa: homogeneous_fast_path = false is not provided by this build");
    if (o.custom_eig || o.custom_multiply || o.custom_solve)
        throw Unsupported("rcwa: custom hooks are not consulted by the GPU solver; use the batched hook kernels");
}
inline void require_gpu_options(const Options& o, bool) {
    require_gpu_options(o);
    if (o.linear != LinearSolver::LUPartial) throw Unsupported("rcwa: this build solves with partial-pivot LU");
}

inline void flat_stack(const Stack& s, std::vector<double>& out) {
    out.insert(out.end(), {s.period, s.nI.real(), s.nI.imag(), s.nII.real(), s.nII.imag(),
                           static_cast<double>(s.layers.size())});
    for (const Layer& l : s.layers) {
        out.push_back(l.d);
        if (l.homogeneous) {
            out.insert(out.end(), {0.0, l.n.real(), l.n.imag()});
        } else {
            out.insert(out.end(), {1.0, static_cast<double>(l.n_values.size())});
            // the record holds exactly K+1 edges; a size mismatch is the reference's invalid_argument
            if (l.edges.size() != l.n_values.size() + 1)
                throw std::invalid_argument("rcwa: a grating layer needs K indices and K+1 edges");
            out.insert(out.end(), l.edges.begin(), l.edges.end());
            for (const cplx& v : l.n_values) out.insert(out.end(), {v.real(), v.imag()});
            out.insert(out.end(), {static_cast<double>(static_cast<int>(l.fourier_method)),
                                   static_cast<double>(l.fft_samples)});
        }
    }
}

inline void flat_stack(const Stack2D& s, std::vector<double>& out) {
    out.insert(out.end(), {s.Lx, s.Ly, s.nI, s.nII.real(), s.nII.imag(), static_cast<double>(s.layers.size())});
    for (const Layer2D& l : s.layers) {
        out.push_back(l.d);
        if (l.homogeneous) {
            out.insert(out.end(), {0.0, l.n.real(), l.n.imag()});
        } else {
            if (l.y_edges.size() != l.x_profiles.size() + 1)
                throw std::invalid_argument("rcwa2d: a patterned layer needs K x-profiles and K+1 y-edges");
            out.insert(out.end(), {1.0, static_cast<double>(l.x_profiles.size())});
            out.insert(out.end(), l.y_edges.begin(), l.y_edges.end());
            for (const XProfile2D& p : l.x_profiles) {
                if (p.edges.size() != p.n_values.size() + 1)
                    throw std::invalid_argument("rcwa2d: an x-profile needs K indices and K+1 edges");
                out.push_back(static_cast<double>(p.n_values.size()));
                out.insert(out.end(), p.edges.begin(), p.edges.end());
                for (const cplx& v : p.n_values) out.insert(out.end(), {v.real(), v.imag()});
            }
        }
    }
}

// Result fields from the plan report.  A batched plan runs every problem at once, so the stage times
// are shared out evenly: timing.eig is the eigensolve stage, timing.linear the modes and S-matrix
// stage (its dense solves and the products between them), timing.product stays 0.
template <class R>
void fill_diagnostics(R& r, const PlanReport& rep, int i, int problems) {
    const double* d = rep.diag.data() + 5 * static_cast<std::size_t>(i);
    r.max_eig_residual = d[0];
    r.max_solve_backward_error = d[1];
    r.n_eig = static_cast<int>(d[2]);
    r.n_dense_solve = static_cast<int>(d[3]);
    const double* f = rep.refine.data() + 3 * static_cast<std::size_t>(i);
    r.refine_fallbacks = static_cast<int>(f[0]);
    r.refine_iterations = static_cast<int>(f[1]);
    r.refine_residual = f[2];
    if (rep.ms[1] >= 0) {
        r.timing.eig = rep.ms[1] * 1e-3 / problems;
        r.timing.linear = rep.ms[2] * 1e-3 / problems;
    }
}

inline std::vector<Result> planar_batch(const Stack& stack, const std::vector<Point>& points, int M, Polarization pol,
                                        const Options& options) {
    require_gpu_options(options, true);
    const auto t0 = Clock::now();
    std::vector<Result> out(points.size());
    if (points.empty()) return out;
    std::vector<double> rec;
    flat_stack(stack, rec);
    const int P = static_cast<int>(points.size());
    const std::int64_t off = 0;
    std::vector<int> idx(P, 0), Ms(P, M), pols(P, pol == Polarization::TE ? 0 : 1);
    std::vector<double> wl(P), th(P);
    for (int i = 0; i < P; ++i) {
        wl[i] = points[i].wavelength;
        th[i] = points[i].theta_deg;
    }
    Plan plan;
    check(srcwa_plan1d_create(context(options.device, plan_flags(options)), 1, rec.data(),
                              &off, P, idx.data(), wl.data(), th.data(), Ms.data(), pols.data(),
                              options.max_workspace_bytes, &plan.p));
    plan.run();
    const PlanReport rep = report(plan.p, P);
    const std::int64_t orders = srcwa_plan_orders(plan.p);
    std::vector<double> R(2 * orders), T(2 * orders), DEr(orders), DEt(orders);
    std::vector<int> status(P), info(P);
    check(srcwa_plan_results(plan.p, R.data(), T.data(), DEr.data(), DEt.data(), status.data(), info.data()));
    const double total = seconds_since(t0) / P;
    const std::size_t N = static_cast<std::size_t>(2 * M + 1);
    for (int i = 0; i < P; ++i) {
        if (status[i] == 2) throw std::runtime_error("rcwa: device workspace exhausted");
        Result& r = out[i];
        const std::size_t a = N * i;
        r.orders.resize(N);
        for (std::size_t j = 0; j < N; ++j) r.orders[j] = static_cast<int>(j) - M;
        r.R = as_complex(R, a, N);
        r.T = as_complex(T, a, N);
        r.DEr.assign(DEr.begin() + a, DEr.begin() + a + N);
        r.DEt.assign(DEt.begin() + a, DEt.begin() + a + N);
        r.eig_info = info[i];
        fill_diagnostics(r, rep, i, P);
        r.timing.total = total;
    }
    return out;
}

// keep: when given, receives the executed plan (a near-field caller keeps its capture on the device);
// extra_flags: plan flags beyond the options' (SRCWA_PLAN_FIELDS).
inline std::vector<ConicalResult> conical_batch(const Stack& stack, const std::vector<ConicalPoint>& points, int M,
                                                const Options& options, int extra_flags = 0,
                                                std::shared_ptr<Plan>* keep = nullptr) {
    require_gpu_options(options, true);
    const auto t0 = Clock::now();
    std::vector<ConicalResult> out(points.size());
    if (points.empty()) return out;
    std::vector<double> rec;
    flat_stack(stack, rec);
    const int P = static_cast<int>(points.size());
    const std::int64_t off = 0;
    std::vector<int> idx(P, 0), Ms(P, M);
    std::vector<double> wl(P), th(P), ph(P);
    for (int i = 0; i < P; ++i) {
        wl[i] = points[i].wavelength;
        th[i] = points[i].theta_deg;
        ph[i] = points[i].phi_deg;
    }
    auto owned = std::make_shared<Plan>();
    Plan& plan = *owned;
    check(srcwa_plan_conical_create(context(options.device, plan_flags(options) | extra_flags), 1,
                                    rec.data(), &off, P, idx.data(), wl.data(), th.data(), ph.data(), Ms.data(),
                                    options.max_workspace_bytes, &plan.p));
    plan.run();
    if (keep) *keep = owned;
    const PlanReport rep = report(plan.p, P);
    const std::int64_t orders = srcwa_plan_orders(plan.p);
    std::vector<double> J(8 * orders), Jt(8 * orders), DEr(4 * orders), DEt(4 * orders), geo(8 * orders);
    std::vector<int> status(P), info(P);
    check(srcwa_plan_results(plan.p, J.data(), Jt.data(), DEr.data(), DEt.data(), status.data(), info.data()));
    check(srcwa_plan_geometry(plan.p, geo.data()));
    const double total = seconds_since(t0) / P;
    const std::size_t N = static_cast<std::size_t>(2 * M + 1);
    const double nI = stack.nI.real();
    for (int i = 0; i < P; ++i) {
        if (status[i] == 2) throw std::runtime_error("rcwa: device workspace exhausted");
        ConicalResult& r = out[i];
        const std::size_t a = N * i;
        r.orders.resize(N);
        r.kx.resize(N);
        r.phi_m_deg.resize(N);
        r.kzI.resize(N);
        r.kzII.resize(N);
        r.flux_r.resize(N);
        r.flux_t.resize(N);
        r.Jr.resize(N);
        r.Jt.resize(N);
        r.DEr.resize(N);
        r.DEt.resize(N);
        r.ky = nI * std::sin(points[i].theta_deg * detail::kPi / 180.0) * std::sin(points[i].phi_deg * detail::kPi / 180.0);
        for (std::size_t j = 0; j < N; ++j) {
            const double* g = geo.data() + 8 * (a + j);
            r.orders[j] = static_cast<int>(j) - M;
            r.kx[j] = g[0];
            r.kzI[j] = cplx{g[1], g[2]};
            r.kzII[j] = cplx{g[3], g[4]};
            r.flux_r[j] = g[5];
            r.flux_t[j] = g[6];
            r.phi_m_deg[j] = g[7];
            const double* jr = J.data() + 8 * (a + j);
            const double* jt = Jt.data() + 8 * (a + j);
            r.Jr[j] = Jones{cplx{jr[0], jr[1]}, cplx{jr[2], jr[3]}, cplx{jr[4], jr[5]}, cplx{jr[6], jr[7]}};
            r.Jt[j] = Jones{cplx{jt[0], jt[1]}, cplx{jt[2], jt[3]}, cplx{jt[4], jt[5]}, cplx{jt[6], jt[7]}};
            const double* er = DEr.data() + 4 * (a + j);
            const double* et = DEt.data() + 4 * (a + j);
            r.DEr[j] = JonesEfficiency{er[0], er[1], er[2], er[3]};
            r.DEt[j] = JonesEfficiency{et[0], et[1], et[2], et[3]};
        }
        r.eig_info = info[i];
        fill_diagnostics(r, rep, i, P);
        r.n_schur = static_cast<int>(rep.diag[5 * i + 4]);
        r.timing.total = total;
    }
    return out;
}

}  // namespace detail

inline Result solve(const Stack& stack, double wavelength, double theta_deg, int M,
                    Polarization pol = Polarization::TE, const Options& options = {}) {
    return detail::planar_batch(stack, {Point{wavelength, theta_deg}}, M, pol, options)[0];
}

inline std::vector<Result> sweep(const Stack& stack, const std::vector<Point>& points, int M,
                                 Polarization pol = Polarization::TE, const Options& options = {}) {
    return detail::planar_batch(stack, points, M, pol, options);
}

inline ConicalResult solve_conical(const Stack& stack, double wavelength, double theta_deg, double phi_deg, int M,
                                   const Options& options = {}) {
    return detail::conical_batch(stack, {ConicalPoint{wavelength, theta_deg, phi_deg}}, M, options)[0];
}

inline std::vector<ConicalResult> sweep_conical(const Stack& stack, const std::vector<ConicalPoint>& points, int M,
                                                const Options& options = {}) {
    return detail::conical_batch(stack, points, M, options);
}

// The reference's near-field capture (rcwa2d.hpp:939-950).  Declared so that reference callers that pass
// the capture argument compile; this build computes crossed-grating near fields nowhere and throws
// Unsupported when one is requested (planar near fields: stablercwa_amd/fields.hpp).
namespace detail2d {
struct FieldCapture2D;
}

inline Result2D solve2d(const Stack2D& stack, double wavelength, double theta_deg, double phi_deg, int M1, int M2,
                        const Jones2D& jones, const Options2D& options = {}, detail2d::FieldCapture2D* capture = nullptr) {
    detail::require_gpu_options(options);
    if (capture) throw Unsupported("rcwa2d: near-field capture is not provided by this build");
    const auto t0 = detail::Clock::now();
    std::vector<double> rec;
    detail::flat_stack(stack, rec);
    const std::int64_t off = 0;
    const int idx = 0, fz = options.factorization == Factorization2D::Li ? 0 : 1;
    const double jv[4] = {jones[0].real(), jones[0].imag(), jones[1].real(), jones[1].imag()};
    detail::Plan plan;
    detail::check(srcwa_plan2d_create(detail::context(options.device, detail::plan_flags(options)),
                                      1, rec.data(), &off, 1, &idx, &wavelength, &theta_deg, &phi_deg, &M1, &M2, jv,
                                      &fz, options.max_workspace_bytes, &plan.p));
    plan.run();
    const detail::PlanReport rep = detail::report(plan.p, 1);
    const std::int64_t N = srcwa_plan_orders(plan.p);
    std::vector<double> R(8 * N), T(8 * N), DEr(N), DEt(N), geo(8 * N);
    int status = 0, info = 0;
    detail::check(srcwa_plan_results(plan.p, R.data(), T.data(), DEr.data(), DEt.data(), &status, &info));
    detail::check(srcwa_plan_geometry(plan.p, geo.data()));
    if (status == 2) throw std::runtime_error("rcwa2d: device workspace exhausted");
    Result2D r;
    r.status = status == 0 ? Status2D::Ok : status == 3 ? Status2D::RayleighAnomaly : Status2D::NumericalFailure;
    if (r.status == Status2D::RayleighAnomaly)
        r.message = "rcwa2d: a diffraction order has kz = 0 exactly in a homogeneous medium (Rayleigh anomaly)";
    else if (r.status == Status2D::NumericalFailure)
        r.message = "rcwa2d: numerical failure (eigen iteration, singular system or non-finite values)";
    r.M1 = M1;
    r.M2 = M2;
    r.N = static_cast<int>(N);
    r.i0 = harmonic_index_2d(0, 0, M1, M2);
    r.k0 = 2.0 * detail::kPi / wavelength;
    r.theta = theta_deg * detail::kPi / 180.0;
    r.phi = phi_deg * detail::kPi / 180.0;
    const std::size_t n = static_cast<std::size_t>(N);
    r.Rs = detail::as_complex(R, 0, n);
    r.Rp = detail::as_complex(R, n, n);
    r.Ex_r = detail::as_complex(R, 2 * n, n);
    r.Ey_r = detail::as_complex(R, 3 * n, n);
    r.Ts = detail::as_complex(T, 0, n);
    r.Tp = detail::as_complex(T, n, n);
    r.Ex_t = detail::as_complex(T, 2 * n, n);
    r.Ey_t = detail::as_complex(T, 3 * n, n);
    r.DEr = DEr;
    r.DEt = DEt;
    r.m.resize(n);
    r.n.resize(n);
    r.Kx.resize(n);
    r.Ky.resize(n);
    r.kzI.resize(n);
    r.kzII.resize(n);
    r.Ez_r.resize(n);
    r.Ez_t.resize(n);
    r.prop_r.resize(n);
    r.prop_t.resize(n);
    for (std::size_t i = 0; i < n; ++i) {
        const double* g = geo.data() + 8 * i;
        r.m[i] = static_cast<int>(i) / (2 * M2 + 1) - M1;
        r.n[i] = static_cast<int>(i) % (2 * M2 + 1) - M2;
        r.Kx[i] = g[0];
        r.Ky[i] = g[1];
        r.kzI[i] = cplx{g[2], g[3]};
        r.kzII[i] = cplx{g[4], g[5]};
        // Ez from the transversality of each plane wave (rcwa2d.hpp:1232-1245)
        const cplx sr = r.Kx[i] * r.Ex_r[i] + r.Ky[i] * r.Ey_r[i];
        const cplx st = r.Kx[i] * r.Ex_t[i] + r.Ky[i] * r.Ey_t[i];
        r.Ez_r[i] = sr / r.kzI[i];
        r.Ez_t[i] = -st / r.kzII[i];
        r.prop_r[i] = r.kzI[i].real() > 1e-12;
        r.prop_t[i] = r.kzII[i].real() > 1e-12;
    }
    r.eig_info = info;
    r.max_eig_residual = rep.diag[0];
    r.n_eig = static_cast<int>(rep.diag[2]);
    r.n_dense_solve = static_cast<int>(rep.diag[3]);
    r.refine_fallbacks = static_cast<int>(rep.refine[0]);
    r.refine_iterations = static_cast<int>(rep.refine[1]);
    r.refine_residual = rep.refine[2];
    if (rep.ms[0] >= 0) {
        r.timing.operators = rep.ms[0] * 1e-3;
        r.timing.eig = rep.ms[1] * 1e-3;
        r.timing.linear = rep.ms[2] * 1e-3;
    }
    r.timing.total = detail::seconds_since(t0);
    return r;
}

}  // namespace rcwa
}  // namespace srcwa_amd
