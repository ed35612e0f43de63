// srcwa_amd/fields.hpp -- planar near fields on the GPU with the reference's public interface:
// rcwa::solve_with_fields, PlanarFieldContext, field_at and field_map (rcwa1d_fields.hpp:14-206) and the
// field types they return (rcwa_fields.hpp:24-180).
//
// The solve runs as a one-problem plan created with SRCWA_PLAN_FIELDS: the S-matrix kernel keeps each
// layer's amplitudes a+ (layer top) and a- (layer bottom), the modes stay in the plan's device workspace,
// and field_at / field_map evaluate them on the device (srcwa_plan1d_field_at / _field_map).  Conical solves
// likewise (solve_conical_with_fields, ConicalFieldContext; rcwa1d_conical_fields.hpp), with the Jones vector
// chosen per call.  A context shares its plan with its copies: the capture is never written after the solve,
// so a copy behaves as the reference's deep copy.  Crossed-grating near fields are not provided by this build.
#pragma once

#include <array>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rcwa.hpp"

namespace srcwa_amd {
namespace rcwa {

enum class FieldComponent { Ex, Ey, Ez, Hx, Hy, Hz };   // H is Z0 H_physical
enum class FieldContextStatus { Empty, Ready, Unavailable };
enum class FieldSolveKind { PlanarTE, PlanarTM, Conical, Periodic2D };
enum class FieldPropagation { None, Exponential, FiniteZero, Schur };

inline std::size_t field_component_index(FieldComponent c) { return static_cast<std::size_t>(c); }

// Column-major complex matrix with the reference Matrix's element access (matrix.hpp: A(i, j) at
// data()[i + j rows()]).
class FieldMatrix {
public:
    FieldMatrix() = default;
    FieldMatrix(int rows, int cols) : rows_(rows), cols_(cols), a_(static_cast<std::size_t>(rows) * cols) {}
    int rows() const noexcept { return rows_; }
    int cols() const noexcept { return cols_; }
    cplx& operator()(int i, int j) { return a_[i + static_cast<std::size_t>(j) * rows_]; }
    const cplx& operator()(int i, int j) const { return a_[i + static_cast<std::size_t>(j) * rows_]; }
    cplx* data() noexcept { return a_.data(); }
    const cplx* data() const noexcept { return a_.data(); }
    bool all_finite() const {
        for (const cplx& v : a_)
            if (!std::isfinite(v.real()) || !std::isfinite(v.imag())) return false;
        return true;
    }

private:
    int rows_ = 0, cols_ = 0;
    std::vector<cplx> a_;
};

struct FieldNormalization {
    double electric_incident_amplitude = 1.0;
    double magnetic_incident_amplitude = 1.0;
};

struct FieldLayerInfo {
    std::size_t original_index = 0;
    int effective_index = -1;
    double z_top = 0.0, z_bottom = 0.0, thickness = 0.0;
    bool homogeneous = true;
    cplx uniform_index = 1.0;
    RealVector edges;
    Vector n_values;
    bool has_modes = false;
    bool propagates = false;
    bool analytic_modes = false;
    int mode_count = 0;
    int finite_zero_mode_count = 0;
    FieldPropagation propagation = FieldPropagation::None;
};

struct FieldMetadata {
    FieldSolveKind kind = FieldSolveKind::PlanarTE;
    double wavelength = 0.0, period = 0.0, k0 = 0.0;
    double theta_deg = 0.0, phi_deg = 0.0, z_bottom = 0.0;
    cplx nI = 1.0, nII = 1.0;
    int truncation_order = 0;
    int incident_basis_columns = 0;
    std::vector<int> orders;
    Vector kxm, kzI, kzII;              // physical wavenumbers; transverse factor exp[-i (kxm x + ky y)]
    double ky = 0.0;
    FieldNormalization normalization;
    std::vector<FieldLayerInfo> layers;
    // crossed-grating fields only (kind Periodic2D); kept so the metadata has the reference's shape
    double period_y = 0.0;
    int truncation_order_y = 0;
    std::vector<int> orders_m, orders_n;
    RealVector kyn;
};

// Each component is [order, depth]; components in the order Ex, Ey, Ez, Hx, Hy, Hz.
struct HarmonicFields {
    std::vector<int> orders;
    RealVector z;
    FieldMetadata metadata;
    std::array<FieldMatrix, 6> components;
    FieldMatrix& component(FieldComponent c) { return components[field_component_index(c)]; }
    const FieldMatrix& component(FieldComponent c) const { return components[field_component_index(c)]; }
    std::size_t nz() const noexcept { return z.size(); }
    std::size_t norders() const noexcept { return orders.size(); }
    cplx& at(FieldComponent c, std::size_t iz, std::size_t order_index) {
        if (iz >= nz() || order_index >= norders()) throw std::out_of_range("rcwa: harmonic field index");
        return component(c)(static_cast<int>(order_index), static_cast<int>(iz));
    }
    const cplx& at(FieldComponent c, std::size_t iz, std::size_t order_index) const {
        if (iz >= nz() || order_index >= norders()) throw std::out_of_range("rcwa: harmonic field index");
        return component(c)(static_cast<int>(order_index), static_cast<int>(iz));
    }
};

// Each component is [x, z]: data()[iz nx + ix]; x_k = k period / nx.
struct FieldGridXZ {
    RealVector x, z;
    double y = 0.0;
    FieldMetadata metadata;
    std::array<FieldMatrix, 6> components;
    FieldMatrix& component(FieldComponent c) { return components[field_component_index(c)]; }
    const FieldMatrix& component(FieldComponent c) const { return components[field_component_index(c)]; }
    std::size_t nx() const noexcept { return x.size(); }
    std::size_t nz() const noexcept { return z.size(); }
    cplx& at(FieldComponent c, std::size_t iz, std::size_t ix) {
        if (iz >= nz() || ix >= nx()) throw std::out_of_range("rcwa: field grid index");
        return component(c)(static_cast<int>(ix), static_cast<int>(iz));
    }
    const cplx& at(FieldComponent c, std::size_t iz, std::size_t ix) const {
        if (iz >= nz() || ix >= nx()) throw std::out_of_range("rcwa: field grid index");
        return component(c)(static_cast<int>(ix), static_cast<int>(iz));
    }
};

struct PlanarFieldResult;
class PlanarFieldContext;
namespace detail {
struct PlanarFieldAccess;
}

class PlanarFieldContext {
public:
    PlanarFieldContext() = default;
    PlanarFieldContext(const PlanarFieldContext&) = default;
    PlanarFieldContext& operator=(const PlanarFieldContext&) = default;
    PlanarFieldContext(PlanarFieldContext&& other) noexcept { *this = std::move(other); }
    PlanarFieldContext& operator=(PlanarFieldContext&& other) noexcept {
        if (this != &other) {
            metadata_ = std::move(other.metadata_);
            plan_ = std::move(other.plan_);
            status_ = std::exchange(other.status_, FieldContextStatus::Empty);
            other.metadata_ = FieldMetadata{};
            other.plan_.reset();
        }
        return *this;
    }
    FieldContextStatus status() const noexcept { return status_; }
    bool available() const noexcept { return status_ == FieldContextStatus::Ready; }
    const FieldMetadata& metadata() const noexcept { return metadata_; }
    void require_available() const {
        if (!available()) throw std::logic_error("rcwa: planar field context is unavailable");
    }

private:
    FieldMetadata metadata_;
    std::shared_ptr<::srcwa_amd::detail::Plan> plan_;
    FieldContextStatus status_ = FieldContextStatus::Empty;
    friend struct detail::PlanarFieldAccess;
    friend PlanarFieldResult solve_with_fields(const Stack&, double, double, int, Polarization, const Options&);
};

struct PlanarFieldResult {
    Result result;
    PlanarFieldContext fields;
};

namespace detail {

// The outer media's kz = sqrt(n^2 - kx^2) on the reference's branch (analytic_modes, rcwa1d.hpp:522-536).
inline cplx outer_kz(cplx n, double kx) {
    cplx kz = std::sqrt(n * n - kx * kx);
    if (kz.imag() < 0.0) kz = -kz;
    if (std::abs(kz.imag()) < 1e-12 && kz.real() < 0.0) kz = -kz;
    return kz;
}

struct PlanarFieldAccess {
    static srcwa_plan* plan(const PlanarFieldContext& c) { return c.plan_ ? c.plan_->p : nullptr; }
};

inline void fill_components(std::array<FieldMatrix, 6>& comps, const std::vector<double>& v, int rows, int cols) {
    const std::size_t n = static_cast<std::size_t>(rows) * cols;
    for (int c = 0; c < 6; ++c) {
        comps[c] = FieldMatrix(rows, cols);
        for (std::size_t k = 0; k < n; ++k) comps[c].data()[k] = cplx{v[2 * (c * n + k)], v[2 * (c * n + k) + 1]};
    }
}

}  // namespace detail

// One solve with its near-field capture (rcwa1d_fields.hpp:55-104).  A failed solve keeps the reference's
// NaN results and leaves the context Unavailable.
inline PlanarFieldResult solve_with_fields(const Stack& stack, double wavelength, double theta_deg, int M,
                                           Polarization pol = Polarization::TE, const Options& options = {}) {
    detail::require_gpu_options(options, true);
    PlanarFieldResult out;
    std::vector<double> rec;
    detail::flat_stack(stack, rec);
    const std::int64_t off = 0;
    int idx = 0, Ms = M, pols = pol == Polarization::TE ? 0 : 1;
    auto plan = std::make_shared<detail::Plan>();
    detail::check(srcwa_plan1d_create(detail::context(options.device, detail::plan_flags(options) | SRCWA_PLAN_FIELDS), 1,
                                  rec.data(), &off, 1, &idx, &wavelength, &theta_deg, &Ms, &pols,
                                  options.max_workspace_bytes, &plan->p));
    plan->run();
    const int N = 2 * M + 1;
    std::vector<double> R(2 * N), T(2 * N), DEr(N), DEt(N);
    int status = 0, info = 0, ok = 0;
    detail::check(srcwa_plan_results(plan->p, R.data(), T.data(), DEr.data(), DEt.data(), &status, &info));
    if (status == 2) throw std::runtime_error("rcwa: device workspace exhausted");
    detail::check(srcwa_plan1d_field_available(plan->p, 0, &ok));
    Result& r = out.result;
    r.orders.resize(N);
    for (int j = 0; j < N; ++j) r.orders[j] = j - M;
    r.R = detail::as_complex(R, 0, N);
    r.T = detail::as_complex(T, 0, N);
    r.DEr = DEr;
    r.DEt = DEt;
    r.eig_info = info;
    const detail::PlanReport rep = detail::report(plan->p, 1);
    detail::fill_diagnostics(r, rep, 0, 1);

    PlanarFieldContext& ctx = out.fields;
    FieldMetadata& meta = ctx.metadata_;
    meta.kind = pol == Polarization::TE ? FieldSolveKind::PlanarTE : FieldSolveKind::PlanarTM;
    meta.wavelength = wavelength;
    meta.k0 = 2.0 * detail::kPi / wavelength;
    meta.theta_deg = theta_deg;
    meta.truncation_order = M;
    meta.incident_basis_columns = 1;
    meta.orders = r.orders;
    meta.normalization.electric_incident_amplitude = pol == Polarization::TE ? 1.0 : 1.0 / stack.nI.real();
    meta.normalization.magnetic_incident_amplitude = pol == Polarization::TE ? stack.nI.real() : 1.0;
    // geometry (copy_field_geometry, rcwa_fields.hpp:403-434, every layer kept)
    meta.period = stack.period;
    meta.nI = stack.nI;
    meta.nII = stack.nII;
    double z = 0.0;
    for (std::size_t j = 0; j < stack.layers.size(); ++j) {
        const Layer& l = stack.layers[j];
        FieldLayerInfo info_j;
        info_j.original_index = j;
        info_j.effective_index = static_cast<int>(j);
        info_j.z_top = z;
        info_j.thickness = l.d;
        z += l.d;
        info_j.z_bottom = z;
        info_j.homogeneous = l.homogeneous;
        info_j.uniform_index = l.n;
        info_j.edges = l.edges;
        info_j.n_values = l.n_values;
        meta.layers.push_back(std::move(info_j));
    }
    meta.z_bottom = z;
    const double s = stack.nI.real() * std::sin(theta_deg * detail::kPi / 180.0);
    bool finite = std::isfinite(meta.k0);
    for (int j = 0; j < N; ++j) {
        const double kx = s - (j - M) * wavelength / stack.period;
        meta.kxm.push_back(meta.k0 * kx);
        meta.kzI.push_back(meta.k0 * detail::outer_kz(stack.nI, kx));
        meta.kzII.push_back(meta.k0 * detail::outer_kz(stack.nII, kx));
        for (const cplx& v : {meta.kxm.back(), meta.kzI.back(), meta.kzII.back()})
            finite = finite && std::isfinite(v.real()) && std::isfinite(v.imag());
    }
    ctx.status_ = ok && finite ? FieldContextStatus::Ready : FieldContextStatus::Unavailable;
    if (ctx.available()) {
        std::vector<int> zeros(meta.layers.size() + 1, 0);
        ::srcwa_amd::detail::check(srcwa_plan1d_field_layers(plan->p, 0, zeros.data()));
        ctx.plan_ = std::move(plan);
        for (std::size_t j = 0; j < meta.layers.size(); ++j) {
            FieldLayerInfo& l = meta.layers[j];
            l.has_modes = true;
            l.propagates = l.thickness > 0.0;
            l.analytic_modes = l.homogeneous;
            l.mode_count = N;
            l.finite_zero_mode_count = zeros[j];
            l.propagation = zeros[j] ? FieldPropagation::FiniteZero : FieldPropagation::Exponential;
        }
    }
    return out;
}

// Harmonic fields at the depths zs (rcwa1d_fields.hpp:193-198): z < 0 the incidence side, z past the
// stack the exit side, an interface belongs to the layer above it.
inline HarmonicFields field_at(const PlanarFieldContext& context, const RealVector& zs) {
    context.require_available();
    for (double v : zs)
        if (!std::isfinite(v)) throw std::invalid_argument("rcwa: field depths must be finite");
    HarmonicFields out;
    out.metadata = context.metadata();
    out.orders = out.metadata.orders;
    out.z = zs;
    const int n = static_cast<int>(out.orders.size()), nz = static_cast<int>(zs.size());
    std::vector<double> v(12 * static_cast<std::size_t>(n) * nz);
    if (nz > 0)
        ::srcwa_amd::detail::check(
            srcwa_plan1d_field_at(detail::PlanarFieldAccess::plan(context), 0, nz, zs.data(), v.data()));
    detail::fill_components(out.components, v, n, nz);
    return out;
}
inline HarmonicFields field_at(const PlanarFieldContext& context, double z) { return field_at(context, RealVector{z}); }
inline HarmonicFields field_at(const PlanarFieldResult& result, const RealVector& zs) { return field_at(result.fields, zs); }
inline HarmonicFields field_at(const PlanarFieldResult& result, double z) { return field_at(result.fields, z); }

// The fields on the grid x_k = k period / nx (k < nx) at height y (rcwa1d_fields.hpp:200-206).
inline FieldGridXZ field_map(const PlanarFieldContext& context, const RealVector& zs, int nx = 256, double y = 0.0) {
    if (nx <= 0) throw std::invalid_argument("rcwa: field nx must be positive");
    if (!std::isfinite(y)) throw std::invalid_argument("rcwa: field y must be finite");
    context.require_available();
    for (double v : zs)
        if (!std::isfinite(v)) throw std::invalid_argument("rcwa: field depths must be finite");
    FieldGridXZ out;
    out.metadata = context.metadata();
    out.z = zs;
    out.y = y;
    out.x.resize(nx);
    for (int ix = 0; ix < nx; ++ix) out.x[ix] = (static_cast<double>(ix) / nx) * out.metadata.period;
    const int nz = static_cast<int>(zs.size());
    std::vector<double> v(12 * static_cast<std::size_t>(nx) * nz);
    if (nz > 0)
        ::srcwa_amd::detail::check(
            srcwa_plan1d_field_map(detail::PlanarFieldAccess::plan(context), 0, nz, zs.data(), nx, y, v.data()));
    detail::fill_components(out.components, v, nx, nz);
    return out;
}
inline FieldGridXZ field_map(const PlanarFieldResult& result, const RealVector& zs, int nx = 256, double y = 0.0) {
    return field_map(result.fields, zs, nx, y);
}

// ---- conical near fields (rcwa1d_conical_fields.hpp:14-262) ----------------------------------------------

using JonesVector = std::array<cplx, 2>;   // [Es, Ep]

// normalized_jones (rcwa_fields.hpp:218-230)
inline JonesVector normalized_jones(JonesVector jones) {
    double scale = 0.0;
    for (cplx v : jones) {
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
            throw std::invalid_argument("rcwa: Jones input must be finite and nonzero");
        scale = (std::max)(scale, (std::max)(std::abs(v.real()), std::abs(v.imag())));
    }
    if (scale == 0.0) throw std::invalid_argument("rcwa: Jones input must be finite and nonzero");
    for (auto& v : jones) v /= scale;
    const double norm = std::sqrt(std::norm(jones[0]) + std::norm(jones[1]));
    for (auto& v : jones) v /= norm;
    return jones;
}

// jones_from_angles (rcwa_fields.hpp:232-237)
inline JonesVector jones_from_angles(double psi_deg = 90.0, double delta_deg = 0.0) {
    if (!std::isfinite(psi_deg) || !std::isfinite(delta_deg))
        throw std::invalid_argument("rcwa: Jones angles must be finite");
    const double psi = psi_deg * ::srcwa_amd::detail::kPi / 180.0, delta = delta_deg * ::srcwa_amd::detail::kPi / 180.0;
    return normalized_jones({std::sin(psi) * std::exp(cplx(0.0, 1.0) * delta), cplx(std::cos(psi), 0.0)});
}

struct ConicalFieldResult;
class ConicalFieldContext;
namespace detail {
struct ConicalFieldAccess {
    static srcwa_plan* plan(const ConicalFieldContext& c);
};
}  // namespace detail

// The capture of one conical solve for both incident polarizations; a Jones vector is chosen per call.
// Degenerate layers are sampled in their Schur basis.
class ConicalFieldContext {
public:
    ConicalFieldContext() = default;
    ConicalFieldContext(const ConicalFieldContext&) = default;
    ConicalFieldContext& operator=(const ConicalFieldContext&) = default;
    ConicalFieldContext(ConicalFieldContext&& other) noexcept { *this = std::move(other); }
    ConicalFieldContext& operator=(ConicalFieldContext&& other) noexcept {
        if (this != &other) {
            metadata_ = std::move(other.metadata_);
            plan_ = std::move(other.plan_);
            status_ = std::exchange(other.status_, FieldContextStatus::Empty);
            other.metadata_ = FieldMetadata{};
            other.plan_.reset();
        }
        return *this;
    }
    FieldContextStatus status() const noexcept { return status_; }
    bool available() const noexcept { return status_ == FieldContextStatus::Ready; }
    const FieldMetadata& metadata() const noexcept { return metadata_; }
    void require_available() const {
        if (!available()) throw std::logic_error("rcwa: conical field context is unavailable");
    }

private:
    FieldMetadata metadata_;
    std::shared_ptr<::srcwa_amd::detail::Plan> plan_;
    FieldContextStatus status_ = FieldContextStatus::Empty;
    friend struct detail::ConicalFieldAccess;
    friend ConicalFieldResult solve_conical_with_fields(const Stack&, double, double, double, int, const Options&);
};

inline srcwa_plan* detail::ConicalFieldAccess::plan(const ConicalFieldContext& c) { return c.plan_ ? c.plan_->p : nullptr; }

struct ConicalFieldResult {
    ConicalResult result;
    ConicalFieldContext fields;
};

// One conical solve with its capture for the s and p incidence (rcwa1d_conical_fields.hpp:56-98).
inline ConicalFieldResult solve_conical_with_fields(const Stack& stack, double wavelength, double theta_deg,
                                                   double phi_deg, int M, const Options& options = {}) {
    ConicalFieldResult out;
    std::shared_ptr<::srcwa_amd::detail::Plan> plan;
    out.result = detail::conical_batch(stack, {ConicalPoint{wavelength, theta_deg, phi_deg}}, M, options,
                                       SRCWA_PLAN_FIELDS, &plan)[0];
    int ok = 0;
    ::srcwa_amd::detail::check(srcwa_plan1d_field_available(plan->p, 0, &ok));
    ConicalFieldContext& ctx = out.fields;
    FieldMetadata& meta = ctx.metadata_;
    meta.kind = FieldSolveKind::Conical;
    meta.wavelength = wavelength;
    meta.k0 = 2.0 * ::srcwa_amd::detail::kPi / wavelength;
    meta.theta_deg = theta_deg;
    meta.phi_deg = phi_deg;
    meta.truncation_order = M;
    meta.incident_basis_columns = 2;
    meta.orders = out.result.orders;
    meta.normalization.electric_incident_amplitude = 1.0;
    meta.normalization.magnetic_incident_amplitude = stack.nI.real();
    // geometry with zero-thickness layers skipped (copy_field_geometry, rcwa_fields.hpp:403-434)
    meta.period = stack.period;
    meta.nI = stack.nI;
    meta.nII = stack.nII;
    double z = 0.0;
    int effective = 0;
    for (std::size_t j = 0; j < stack.layers.size(); ++j) {
        const Layer& l = stack.layers[j];
        FieldLayerInfo info_j;
        info_j.original_index = j;
        info_j.effective_index = l.d == 0.0 ? -1 : effective++;
        info_j.z_top = z;
        info_j.thickness = l.d;
        z += l.d;
        info_j.z_bottom = z;
        info_j.homogeneous = l.homogeneous;
        info_j.uniform_index = l.n;
        info_j.edges = l.edges;
        info_j.n_values = l.n_values;
        meta.layers.push_back(std::move(info_j));
    }
    meta.z_bottom = z;
    meta.ky = meta.k0 * out.result.ky;
    bool finite = std::isfinite(meta.k0) && std::isfinite(meta.ky);
    for (std::size_t j = 0; j < out.result.kx.size(); ++j) {
        meta.kxm.push_back(meta.k0 * out.result.kx[j]);
        meta.kzI.push_back(meta.k0 * out.result.kzI[j]);
        meta.kzII.push_back(meta.k0 * out.result.kzII[j]);
        for (const cplx& v : {meta.kxm.back(), meta.kzI.back(), meta.kzII.back()})
            finite = finite && std::isfinite(v.real()) && std::isfinite(v.imag());
    }
    ctx.status_ = ok && finite ? FieldContextStatus::Ready : FieldContextStatus::Unavailable;
    if (ctx.available()) {
        std::vector<int> schur(static_cast<std::size_t>(effective) + 1, 0);
        ::srcwa_amd::detail::check(srcwa_plan_conical_field_layers(plan->p, 0, schur.data()));
        ctx.plan_ = std::move(plan);
        const int modes = 2 * static_cast<int>(out.result.orders.size());
        for (FieldLayerInfo& l : meta.layers) {
            if (l.effective_index < 0) continue;
            l.has_modes = true;
            l.propagates = true;
            l.analytic_modes = l.homogeneous;
            l.mode_count = modes;
            l.propagation = schur[l.effective_index] ? FieldPropagation::Schur : FieldPropagation::Exponential;
        }
    }
    return out;
}

inline HarmonicFields field_at(const ConicalFieldContext& context, const RealVector& zs,
                               JonesVector jones = {{1.0, 0.0}}) {
    context.require_available();
    jones = normalized_jones(jones);
    for (double v : zs)
        if (!std::isfinite(v)) throw std::invalid_argument("rcwa: field depths must be finite");
    HarmonicFields out;
    out.metadata = context.metadata();
    out.orders = out.metadata.orders;
    out.z = zs;
    const int n = static_cast<int>(out.orders.size()), nz = static_cast<int>(zs.size());
    const double jv[4] = {jones[0].real(), jones[0].imag(), jones[1].real(), jones[1].imag()};
    std::vector<double> v(12 * static_cast<std::size_t>(n) * nz);
    if (nz > 0)
        ::srcwa_amd::detail::check(
            srcwa_plan_conical_field_at(detail::ConicalFieldAccess::plan(context), 0, nz, zs.data(), jv, v.data()));
    detail::fill_components(out.components, v, n, nz);
    return out;
}
inline HarmonicFields field_at(const ConicalFieldContext& context, double z, JonesVector jones = {{1.0, 0.0}}) {
    return field_at(context, RealVector{z}, jones);
}
inline HarmonicFields field_at(const ConicalFieldResult& result, const RealVector& zs, JonesVector jones = {{1.0, 0.0}}) {
    return field_at(result.fields, zs, jones);
}
inline HarmonicFields field_at(const ConicalFieldResult& result, double z, JonesVector jones = {{1.0, 0.0}}) {
    return field_at(result.fields, z, jones);
}

inline FieldGridXZ field_map(const ConicalFieldContext& context, const RealVector& zs, int nx = 256, double y = 0.0,
                             JonesVector jones = {{1.0, 0.0}}) {
    if (nx <= 0) throw std::invalid_argument("rcwa: field nx must be positive");
    if (!std::isfinite(y)) throw std::invalid_argument("rcwa: field y must be finite");
    context.require_available();
    jones = normalized_jones(jones);
    for (double v : zs)
        if (!std::isfinite(v)) throw std::invalid_argument("rcwa: field depths must be finite");
    FieldGridXZ out;
    out.metadata = context.metadata();
    out.z = zs;
    out.y = y;
    out.x.resize(nx);
    for (int ix = 0; ix < nx; ++ix) out.x[ix] = (static_cast<double>(ix) / nx) * out.metadata.period;
    const int nz = static_cast<int>(zs.size());
    const double jv[4] = {jones[0].real(), jones[0].imag(), jones[1].real(), jones[1].imag()};
    std::vector<double> v(12 * static_cast<std::size_t>(nx) * nz);
    if (nz > 0)
        ::srcwa_amd::detail::check(srcwa_plan_conical_field_map(detail::ConicalFieldAccess::plan(context), 0, nz,
                                                                zs.data(), nx, y, jv, v.data()));
    detail::fill_components(out.components, v, nx, nz);
    return out;
}
inline FieldGridXZ field_map(const ConicalFieldResult& result, const RealVector& zs, int nx = 256, double y = 0.0,
                             JonesVector jones = {{1.0, 0.0}}) {
    return field_map(result.fields, zs, nx, y, jones);
}

}  // namespace rcwa
}  // namespace srcwa_amd
