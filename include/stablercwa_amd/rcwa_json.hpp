// JSON request / reply layer over the GPU solver: the request is one JSON text, the reply is one JSON text.
// Replaces the reference's rcwa::solve_json / solve_json_reply (include/stablercwa/rcwa_json.hpp:1287-1312)
// with the same request keys, the same reply layout (members in the same order) and the same error
// convention; only the wording of error messages is this build's own (English).
//
//   std::string reply = srcwa_amd::rcwa::solve_json(R"({"mode": "planar",
//       "stack": {"nI": 1.0, "nII": 1.46, "period": 0.574,
//                 "layers": [{"d": 0.38, "edges": [0.0, 0.172, 0.574], "n_values": [1.46, 1.0]},
//                            {"d": 0.115, "n": [1.96, 0.0]}]},
//       "M": 20, "polarization": "TE", "wavelength": 1.053, "theta_deg": 66.36})");
//
// Request (members in any order; an unknown key is an error):
//   mode          "planar" (TE / TM in the plane of incidence), "conical" (any azimuth; Jones matrices per
//                 order) or "periodic2d" (crossed gratings)
//   stack         planar / conical: {nI (positive real), nII, period, layers: [{d, n} | {d, edges, n_values,
//                 fourier_method ("analytic" | "fft"), fft_samples}]}; periodic2d: {nI, nII, Lx, Ly,
//                 layers: [{d, n} | {d, y_edges, x_profiles: [{edges, n_values}]}]}.  An index is a real
//                 number or [re, im]
//   M             truncation order (orders -M..M); periodic2d also My
//   polarization  planar only: "TE" or "TM"
//   points        [{wavelength, theta_deg, phi_deg (conical, periodic2d)}], or wavelength / theta_deg / phi_deg
//                 at the top level for a single point
//   input         conical (optional) and periodic2d (required): {psi_deg, delta_deg} or {jones: [Es, Ep]}
//   options       backend ("builtin"), eig ("qr"), linear ("lu", planar and conical), factorization ("li" |
//                 "laurent", periodic2d), homogeneous_fast_path (true), parallel_layers, diagnostics, threads
//                 (1..256; accepted and echoed: the solves run on the GPU).  The reference's "mkl" backend,
//                 "lr" eigensolver, "lu_nopivot" / "qr" solvers and homogeneous_fast_path = false are not
//                 provided by this build: such a request is answered with ok = false at that option's path.
//
// Reply (members in this order):
//   {"ok": true, "version": 1, "mode": ..., "M": M, ["My": My,] "orders": [...], "threads": t, "results": [...]}
//   planar result: wavelength, theta_deg, R, T ([re, im] per order), DEr, DEt, R_total, T_total, eig_info,
//     n_eig, n_dense_solve, max_eig_residual, max_solve_backward_error, timing {total, eig, linear, product}
//   conical result: wavelength, theta_deg, phi_deg, kx, ky, phi_m_deg, kzI, kzII, flux_r, flux_t, Jr, Jt
//     ([ss, sp, ps, pp] per order), DEr_channels, DEt_channels, flux_unitarity, polarized (with input:
//     jones_in, Rs, Rp, Ts, Tp, DEr, DEt, DEr_s, DEr_p, DEt_s, DEt_p, R_total, T_total), the bookkeeping
//     members of planar, n_schur, min_abs_lambda (null when +inf), degenerate_mode_warning
//   periodic2d result: wavelength, theta_deg, phi_deg, status ("ok" | "rayleigh_anomaly" |
//     "numerical_failure"), message, Kx, Ky, kzI, kzII, Rs, Rp, Ts, Tp, Ex_r, Ey_r, Ez_r, Ex_t, Ey_t, Ez_t, DEr,
//     DEt (all null when status is not "ok"), prop_r, prop_t (0 / 1), eig_info, n_eig, n_dense_solve,
//     max_eig_residual, max_solve_backward_error (0), timing {total, operators, eig, linear, product}
//   A numerical failure is not a request error: ok stays true and the failed numbers are null.
//   A request error: {"ok": false, "error": "...", "where": "stack.layers[1].n_values[0]"} (where is empty when
//   the text is not JSON or its top level is not an object).  Numbers read back bitwise as the solver
//   returned them.  solve_json never throws.
#pragma once

#include <array>
#include <cmath>
#include <exception>
#include <initializer_list>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "json.hpp"
#include "rcwa.hpp"

namespace srcwa_amd {
namespace rcwa {

constexpr int kJsonVersion = 1;
constexpr int kJsonMaxThreads = 256;
// crossed gratings: (2M+1)(2My+1) harmonics at most (a 4096-order dense matrix), checked before anything
// is allocated; planar and conical: the dense order (2M+1, or 2(2M+1) for conical) at most 16384
constexpr int kJsonMax2dHarmonics = 2048;
constexpr int kJsonMaxMatrixOrder = 16384;

struct JsonReply {
    bool ok = false;
    std::string text;
};

namespace json_detail {

using srcwa_amd::json::Value;
using srcwa_amd::json::Writer;

struct RequestError {
    std::string message, where;
};
[[noreturn]] inline void fail(const std::string& where, const std::string& message) { throw RequestError{message, where}; }

inline std::string child(const std::string& path, const std::string& key) { return path.empty() ? key : path + "." + key; }
inline std::string child(const std::string& path, std::size_t i) { return path + "[" + std::to_string(i) + "]"; }

inline void only_keys(const Value& obj, const std::string& path, std::initializer_list<const char*> allowed) {
    for (std::size_t i = 0; i < obj.size(); ++i) {
        const std::string& k = obj.key(i);
        bool ok = false;
        for (const char* a : allowed) ok = ok || k == a;
        if (ok) continue;
        std::string list;
        for (const char* a : allowed) list += (list.empty() ? "" : ", ") + std::string(a);
        const std::string shown = k.size() <= 48 ? k : k.substr(0, 48) + "...";
        fail(child(path, shown), "unknown key (allowed here: " + list + ")");
    }
}
inline const Value& required(const Value& obj, const char* key, const std::string& path) {
    const Value* v = obj.find(key);
    if (!v) fail(child(path, key), "missing");
    return *v;
}
inline void want_object(const Value& v, const std::string& where) {
    if (!v.is_object()) fail(where, "must be an object {...}");
}
inline void want_array(const Value& v, const std::string& where) {
    if (!v.is_array()) fail(where, "must be an array [...]");
}
inline double get_number(const Value& v, const std::string& where) {
    if (!v.is_number()) fail(where, "must be a number");
    return v.as_double();
}
inline int get_int(const Value& v, const std::string& where) {
    int out = 0;
    if (!v.to_int(out)) fail(where, "must be an integer");
    return out;
}
inline bool get_bool(const Value& v, const std::string& where) {
    if (!v.is_bool()) fail(where, "must be true or false");
    return v.as_bool();
}
inline std::string get_text(const Value& v, const std::string& where) {
    if (!v.is_string()) fail(where, "must be a string");
    return v.as_string();
}
inline cplx get_complex(const Value& v, const std::string& where) {
    if (v.is_number()) return cplx{v.as_double(), 0.0};
    if (!v.is_array() || v.size() != 2 || !v.at(0).is_number() || !v.at(1).is_number())
        fail(where, "must be a real number or [re, im]");
    return cplx{v.at(0).as_double(), v.at(1).as_double()};
}

// The solver's index rule (rcwa1d.hpp:201-204): finite, nonzero, Re >= 0, Im >= 0.
inline void index_ok(cplx n, const std::string& where) {
    if (!std::isfinite(n.real()) || !std::isfinite(n.imag()) || std::abs(n) == 0.0 || n.real() < 0.0 || n.imag() < 0.0)
        fail(where, "a refractive index must be finite and nonzero with non-negative real and imaginary parts "
                    "(absorbing media as [n, k], k > 0)");
}

struct PointSpec {
    double wavelength = 1.0, theta_deg = 0.0, phi_deg = 0.0;
};
enum class InputKind { None, Angles, Jones };

struct Request {
    std::string mode = "planar";
    Stack stack;
    Stack2D stack2d;
    int M = 0, My = 0;
    Polarization polarization = Polarization::TE;
    std::vector<PointSpec> points;
    Options options;
    Options2D options2d;
    int threads = 0;
    InputKind input = InputKind::None;
    double psi_deg = 0.0, delta_deg = 0.0;
    cplx Es, Ep;
    bool conical() const { return mode == "conical"; }
    bool periodic2d() const { return mode == "periodic2d"; }
};

// increasing finite edges spanning one period to tol * period
inline void edges_ok(const std::vector<double>& e, const std::string& path, double period, double tol,
                     const std::string& period_name) {
    for (std::size_t j = 0; j < e.size(); ++j)
        if (!std::isfinite(e[j]) || (j > 0 && e[j] <= e[j - 1])) fail(child(path, j), "edges must increase strictly");
    if (std::abs((e.back() - e.front()) - period) > tol * period)
        fail(path, "the first and last edge must be exactly one period (" + period_name + ") apart");
}
inline std::vector<double> numbers(const Value& v, const std::string& path) {
    want_array(v, path);
    std::vector<double> out(v.size());
    for (std::size_t j = 0; j < out.size(); ++j) out[j] = get_number(v.at(j), child(path, j));
    return out;
}
inline std::vector<cplx> indices(const Value& v, const std::string& path) {
    want_array(v, path);
    std::vector<cplx> out(v.size());
    for (std::size_t j = 0; j < out.size(); ++j) {
        out[j] = get_complex(v.at(j), child(path, j));
        index_ok(out[j], child(path, j));
    }
    return out;
}

inline Layer read_layer(const Value& v, const std::string& path, double period) {
    want_object(v, path);
    only_keys(v, path, {"d", "n", "edges", "n_values", "fourier_method", "fft_samples"});
    const double d = get_number(required(v, "d", path), child(path, "d"));
    if (!std::isfinite(d) || d < 0.0) fail(child(path, "d"), "a layer thickness cannot be negative");
    const Value* n = v.find("n");
    const Value* edges = v.find("edges");
    const Value* values = v.find("n_values");
    if (n && (edges || values)) fail(path, "a uniform layer has n, a grating layer edges and n_values; not both");
    if (!n && !edges && !values) fail(path, "no index: a uniform layer needs n, a grating layer edges and n_values");
    Layer layer;
    if (n) {
        const cplx idx = get_complex(*n, child(path, "n"));
        index_ok(idx, child(path, "n"));
        layer = Layer::uniform(d, idx);
    } else {
        const std::string ep = child(path, "edges"), np = child(path, "n_values");
        if (!edges) fail(ep, "a grating layer needs its segment edges");
        if (!values) fail(np, "a grating layer needs the index of each segment");
        std::vector<double> e = numbers(*edges, ep);
        std::vector<cplx> idx = indices(*values, np);
        if (idx.empty() || e.size() != idx.size() + 1)
            fail(np, "K segments need K+1 edges (got " + std::to_string(idx.size()) + " indices and " +
                         std::to_string(e.size()) + " edges)");
        edges_ok(e, ep, period, 1e-10, "stack.period");
        layer = Layer::grating(d, std::move(e), std::move(idx));
    }
    if (const Value* m = v.find("fourier_method")) {
        const std::string w = child(path, "fourier_method");
        const std::string name = get_text(*m, w);
        if (name == "analytic") layer.fourier_method = FourierMethod::Analytic;
        else if (name == "fft") layer.fourier_method = FourierMethod::FFT;
        else fail(w, "must be \"analytic\" or \"fft\"");
    }
    if (const Value* s = v.find("fft_samples")) {
        const std::string w = child(path, "fft_samples");
        layer.fft_samples = get_int(*s, w);
        if (layer.fft_samples < 0 || layer.fft_samples > 4194304)
            fail(w, "the FFT sample count must lie in 0..4194304 (0 chooses it)");
    }
    return layer;
}

inline cplx read_incidence_index(const Value& v, const std::string& where) {
    const cplx nI = get_complex(v, where);
    index_ok(nI, where);
    if (nI.imag() != 0.0 || nI.real() <= 0.0) fail(where, "the incidence medium must be lossless (a positive real index)");
    return nI;
}

inline Stack read_stack(const Value& v, bool conical) {
    const std::string path = "stack";
    want_object(v, path);
    only_keys(v, path, {"nI", "nII", "period", "layers"});
    Stack s;
    s.nI = read_incidence_index(required(v, "nI", path), "stack.nI");
    s.nII = get_complex(required(v, "nII", path), "stack.nII");
    index_ok(s.nII, "stack.nII");
    if (conical && s.nII.imag() != 0.0) fail("stack.nII", "conical incidence needs a real index in the exit medium");
    s.period = get_number(required(v, "period", path), "stack.period");
    if (!std::isfinite(s.period) || s.period <= 0.0) fail("stack.period", "the period must be positive");
    const Value& layers = required(v, "layers", path);
    want_array(layers, "stack.layers");
    for (std::size_t i = 0; i < layers.size(); ++i)
        s.layers.push_back(read_layer(layers.at(i), child("stack.layers", i), s.period));
    return s;
}

inline Layer2D read_layer_2d(const Value& v, const std::string& path, double Lx, double Ly) {
    want_object(v, path);
    only_keys(v, path, {"d", "n", "y_edges", "x_profiles"});
    const double d = get_number(required(v, "d", path), child(path, "d"));
    if (!std::isfinite(d) || d < 0.0) fail(child(path, "d"), "a layer thickness cannot be negative");
    const Value* n = v.find("n");
    const Value* ye = v.find("y_edges");
    const Value* xp = v.find("x_profiles");
    if (n && (ye || xp)) fail(path, "a uniform layer has n, a patterned layer y_edges and x_profiles; not both");
    if (!n && !ye && !xp) fail(path, "no index: a uniform layer needs n, a patterned layer y_edges and x_profiles");
    if (n) {
        const cplx idx = get_complex(*n, child(path, "n"));
        index_ok(idx, child(path, "n"));
        return Layer2D::uniform(d, idx);
    }
    const std::string yp = child(path, "y_edges"), pp = child(path, "x_profiles");
    if (!ye) fail(yp, "a patterned layer needs its y edges");
    if (!xp) fail(pp, "a patterned layer needs one x profile per y segment");
    std::vector<double> y = numbers(*ye, yp);
    want_array(*xp, pp);
    if (xp->size() == 0 || y.size() != xp->size() + 1)
        fail(pp, "K x profiles need K+1 y edges (got " + std::to_string(xp->size()) + " profiles and " +
                     std::to_string(y.size()) + " edges)");
    edges_ok(y, yp, Ly, 1e-9, "stack.Ly");
    std::vector<XProfile2D> rows(xp->size());
    for (std::size_t k = 0; k < rows.size(); ++k) {
        const std::string rp = child(pp, k);
        const Value& item = xp->at(k);
        want_object(item, rp);
        only_keys(item, rp, {"edges", "n_values"});
        const std::string ep = child(rp, "edges"), np = child(rp, "n_values");
        rows[k].edges = numbers(required(item, "edges", rp), ep);
        rows[k].n_values = indices(required(item, "n_values", rp), np);
        if (rows[k].n_values.empty() || rows[k].edges.size() != rows[k].n_values.size() + 1)
            fail(np, "K segments need K+1 edges (got " + std::to_string(rows[k].n_values.size()) + " indices and " +
                         std::to_string(rows[k].edges.size()) + " edges)");
        edges_ok(rows[k].edges, ep, Lx, 1e-9, "stack.Lx");
    }
    return Layer2D::patterned(d, std::move(y), std::move(rows));
}

inline Stack2D read_stack_2d(const Value& v) {
    const std::string path = "stack";
    want_object(v, path);
    if (v.contains("period")) fail("stack.period", "periodic2d takes Lx and Ly instead of period");
    only_keys(v, path, {"nI", "nII", "Lx", "Ly", "layers"});
    Stack2D s;
    s.nI = read_incidence_index(required(v, "nI", path), "stack.nI").real();
    s.nII = get_complex(required(v, "nII", path), "stack.nII");
    index_ok(s.nII, "stack.nII");
    s.Lx = get_number(required(v, "Lx", path), "stack.Lx");
    if (!std::isfinite(s.Lx) || s.Lx <= 0.0) fail("stack.Lx", "the period must be positive");
    s.Ly = get_number(required(v, "Ly", path), "stack.Ly");
    if (!std::isfinite(s.Ly) || s.Ly <= 0.0) fail("stack.Ly", "the period must be positive");
    const Value& layers = required(v, "layers", path);
    want_array(layers, "stack.layers");
    for (std::size_t i = 0; i < layers.size(); ++i)
        s.layers.push_back(read_layer_2d(layers.at(i), child("stack.layers", i), s.Lx, s.Ly));
    return s;
}

inline PointSpec read_point(const Value& v, const std::string& path, bool azimuth) {
    PointSpec p;
    p.wavelength = get_number(required(v, "wavelength", path), child(path, "wavelength"));
    if (!std::isfinite(p.wavelength) || p.wavelength <= 0.0) fail(child(path, "wavelength"), "the wavelength must be positive");
    p.theta_deg = get_number(required(v, "theta_deg", path), child(path, "theta_deg"));
    if (!std::isfinite(p.theta_deg) || std::abs(p.theta_deg) >= 90.0)
        fail(child(path, "theta_deg"), "the incidence angle must lie strictly between -90 and 90 degrees");
    if (azimuth)
        if (const Value* phi = v.find("phi_deg")) {
            p.phi_deg = get_number(*phi, child(path, "phi_deg"));
            if (!std::isfinite(p.phi_deg)) fail(child(path, "phi_deg"), "the azimuth must be finite");
        }
    return p;
}

inline const char* no_azimuth() { return "planar incidence has no azimuth; use mode \"conical\" for phi != 0"; }

inline void read_points(const Value& root, Request& rq) {
    const Value* pts = root.find("points");
    const bool single = root.contains("wavelength") || root.contains("theta_deg") || root.contains("phi_deg");
    if (pts && single) fail("points", "give either a points array or a top-level wavelength / theta_deg, not both");
    if (!pts && !single) fail("points", "no incidence: give a points array or a top-level wavelength and theta_deg");
    const bool azimuth = !rq.mode.empty() && rq.mode != "planar";
    if (!pts) {
        rq.points.push_back(read_point(root, "", azimuth));
        return;
    }
    want_array(*pts, "points");
    if (pts->size() == 0) fail("points", "at least one point is needed");
    for (std::size_t k = 0; k < pts->size(); ++k) {
        const std::string path = child("points", k);
        const Value& item = pts->at(k);
        want_object(item, path);
        if (azimuth) {
            only_keys(item, path, {"wavelength", "theta_deg", "phi_deg"});
        } else {
            if (item.contains("phi_deg")) fail(child(path, "phi_deg"), no_azimuth());
            only_keys(item, path, {"wavelength", "theta_deg"});
        }
        rq.points.push_back(read_point(item, path, azimuth));
    }
}

// options this build provides (the rest are refused at their path, never silently replaced)
inline void read_common_options(const Value& v, Request& rq, bool& diagnostics, bool& parallel_layers) {
    if (const Value* b = v.find("backend")) {
        const std::string name = get_text(*b, "options.backend");
        if (name == "mkl") fail("options.backend", "this build solves on the GPU; the oneMKL backend is not provided");
        if (name != "builtin") fail("options.backend", "must be \"builtin\" (or the reference's \"mkl\")");
    }
    if (const Value* e = v.find("eig")) {
        const std::string name = get_text(*e, "options.eig");
        if (name == "lr") fail("options.eig", "this build's eigensolver is the shifted QR; \"lr\" is not provided");
        if (name != "qr") fail("options.eig", "must be \"qr\" (or the reference's \"lr\")");
    }
    if (const Value* f = v.find("homogeneous_fast_path"))
        if (!get_bool(*f, "options.homogeneous_fast_path"))
            fail("options.homogeneous_fast_path", "homogeneous_fast_path = false is not provided by this build");
    if (const Value* p = v.find("parallel_layers")) parallel_layers = get_bool(*p, "options.parallel_layers");
    if (const Value* d = v.find("diagnostics")) diagnostics = get_bool(*d, "options.diagnostics");
    if (const Value* t = v.find("threads")) {
        rq.threads = get_int(*t, "options.threads");
        if (rq.threads < 1 || rq.threads > kJsonMaxThreads)
            fail("options.threads", "the thread count must lie in 1.." + std::to_string(kJsonMaxThreads));
    }
}

inline void read_options(const Value& v, Request& rq) {
    want_object(v, "options");
    only_keys(v, "options", {"backend", "eig", "linear", "homogeneous_fast_path", "parallel_layers", "diagnostics", "threads"});
    read_common_options(v, rq, rq.options.diagnostics, rq.options.parallel_layers);
    if (const Value* l = v.find("linear")) {
        const std::string name = get_text(*l, "options.linear");
        if (name == "lu_nopivot" || name == "qr")
            fail("options.linear", "this build solves with partial-pivot LU; \"" + name + "\" is not provided");
        if (name != "lu") fail("options.linear", "must be \"lu\" (or the reference's \"lu_nopivot\", \"qr\")");
    }
}

inline void read_options_2d(const Value& v, Request& rq) {
    want_object(v, "options");
    if (v.contains("linear")) fail("options.linear", "crossed-grating solves have no linear-solver option");
    only_keys(v, "options", {"backend", "eig", "factorization", "homogeneous_fast_path", "parallel_layers", "diagnostics",
                             "threads"});
    read_common_options(v, rq, rq.options2d.diagnostics, rq.options2d.parallel_layers);
    if (const Value* f = v.find("factorization")) {
        const std::string name = get_text(*f, "options.factorization");
        if (name == "li") rq.options2d.factorization = Factorization2D::Li;
        else if (name == "laurent") rq.options2d.factorization = Factorization2D::Laurent;
        else fail("options.factorization", "must be \"li\" or \"laurent\"");
    }
}

inline void read_input(const Value& v, Request& rq) {
    want_object(v, "input");
    only_keys(v, "input", {"psi_deg", "delta_deg", "jones"});
    if (const Value* j = v.find("jones")) {
        if (v.contains("psi_deg") || v.contains("delta_deg"))
            fail("input", "give the input polarization either as psi_deg (and delta_deg) or as jones");
        if (!j->is_array() || j->size() != 2) fail("input.jones", "must be the two components [Es, Ep] (s first)");
        rq.Es = get_complex(j->at(0), "input.jones[0]");
        rq.Ep = get_complex(j->at(1), "input.jones[1]");
        const double len = std::hypot(std::abs(rq.Es), std::abs(rq.Ep));
        if (!(len > 0.0) || !std::isfinite(len)) fail("input.jones", "the input Jones vector cannot be zero");
        rq.input = InputKind::Jones;
        return;
    }
    if (!v.contains("psi_deg")) fail("input.psi_deg", "missing (give psi_deg, or jones instead)");
    rq.psi_deg = get_number(v.at("psi_deg"), "input.psi_deg");
    if (const Value* d = v.find("delta_deg")) rq.delta_deg = get_number(*d, "input.delta_deg");
    if (!std::isfinite(rq.psi_deg)) fail("input.psi_deg", "must be finite");
    if (!std::isfinite(rq.delta_deg)) fail("input.delta_deg", "must be finite");
    rq.input = InputKind::Angles;
}

inline int truncation_2d(const Value& root, const char* name) {
    const int m = get_int(required(root, name, ""), name);
    const int limit = (kJsonMax2dHarmonics - 1) / 2;
    if (m < 0 || m > limit)
        fail(name, "the truncation order must lie in 0.." + std::to_string(limit) + " ((2M+1)(2My+1) <= " +
                       std::to_string(kJsonMax2dHarmonics) + ")");
    return m;
}

inline Request read_request(const Value& root) {
    if (!root.is_object()) fail("", "the request must be a JSON object {...}");
    Request rq;
    rq.mode = get_text(required(root, "mode", ""), "mode");
    if (rq.mode != "planar" && rq.mode != "conical" && rq.mode != "periodic2d")
        fail("mode", "must be \"planar\", \"conical\" or \"periodic2d\"");
    if (rq.periodic2d()) {
        if (root.contains("polarization")) fail("polarization", "periodic2d takes its input polarization in input");
        if (root.contains("period")) fail("period", "periodic2d takes stack.Lx and stack.Ly instead of period");
        only_keys(root, "", {"mode", "stack", "M", "My", "points", "wavelength", "theta_deg", "phi_deg", "input", "options"});
        rq.stack2d = read_stack_2d(required(root, "stack", ""));
        rq.M = truncation_2d(root, "M");
        rq.My = truncation_2d(root, "My");
        const long long harmonics = (2LL * rq.M + 1) * (2LL * rq.My + 1);
        if (harmonics > kJsonMax2dHarmonics)
            fail("My", "(2M+1)(2My+1) = " + std::to_string(harmonics) + " harmonics exceed " +
                           std::to_string(kJsonMax2dHarmonics));
        read_points(root, rq);
        const Value* in = root.find("input");
        if (!in) fail("input", "missing: crossed gratings need the input polarization (psi_deg and delta_deg, or jones)");
        read_input(*in, rq);
        if (const Value* o = root.find("options")) read_options_2d(*o, rq);
        return rq;
    }
    if (rq.conical()) {
        if (root.contains("polarization")) fail("polarization", "conical incidence takes its input polarization in input");
        only_keys(root, "", {"mode", "stack", "M", "points", "wavelength", "theta_deg", "phi_deg", "input", "options"});
    } else {
        if (root.contains("phi_deg")) fail("phi_deg", no_azimuth());
        if (root.contains("input")) fail("input", "planar incidence takes its polarization in polarization");
        only_keys(root, "", {"mode", "stack", "M", "polarization", "points", "wavelength", "theta_deg", "options"});
    }
    rq.stack = read_stack(required(root, "stack", ""), rq.conical());
    rq.M = get_int(required(root, "M", ""), "M");
    const int limit = (kJsonMaxMatrixOrder / (rq.conical() ? 2 : 1) - 1) / 2;
    if (rq.M < 0 || rq.M > limit) fail("M", "the truncation order must lie in 0.." + std::to_string(limit));
    if (!rq.conical()) {
        const std::string p = get_text(required(root, "polarization", ""), "polarization");
        if (p == "TE") rq.polarization = Polarization::TE;
        else if (p == "TM") rq.polarization = Polarization::TM;
        else fail("polarization", "must be \"TE\" or \"TM\"");
    }
    read_points(root, rq);
    if (rq.conical())
        if (const Value* in = root.find("input")) read_input(*in, rq);
    if (const Value* o = root.find("options")) read_options(*o, rq);
    return rq;
}

// The solver's own refusals after the request checks: an invalid_argument (input the checks above let
// through) and this build's Unsupported (for instance a truncation order above its limit).
template <class F>
auto run_solver(F&& f, const std::string& order_path) -> decltype(f()) {
    try {
        return f();
    } catch (const std::invalid_argument& e) {
        fail("", std::string("the solver rejected the request: ") + e.what());
    } catch (const Unsupported& e) {
        fail(order_path, std::string("not provided by this build: ") + e.what());
    } catch (const std::bad_alloc&) {
        fail(order_path, "out of memory: the truncation order is too large");
    }
}

// ---- reply ---------------------------------------------------------------------------------------
inline void put_complex(Writer& w, cplx z) {
    w.begin_array(true);
    w.value(z.real());
    w.value(z.imag());
    w.end_array();
}
inline void put_complexes(Writer& w, const std::vector<cplx>& v) {
    w.begin_array(true);
    for (const cplx& z : v) put_complex(w, z);
    w.end_array();
}
inline void put_reals(Writer& w, const std::vector<double>& v) {
    w.begin_array(true);
    for (double x : v) w.value(x);
    w.end_array();
}

inline void head(Writer& w, const Request& rq) {
    w.begin_object();
    w.key("ok");
    w.value(true);
    w.key("version");
    w.value(kJsonVersion);
    w.key("mode");
    w.value(rq.mode);
    w.key("M");
    w.value(rq.M);
    if (rq.periodic2d()) {
        w.key("My");
        w.value(rq.My);
    }
    w.key("orders");
    w.begin_array(true);
    for (int m = -rq.M; m <= rq.M; ++m) {
        if (!rq.periodic2d()) {
            w.value(m);
            continue;
        }
        for (int n = -rq.My; n <= rq.My; ++n) {
            w.begin_array(true);
            w.value(m);
            w.value(n);
            w.end_array();
        }
    }
    w.end_array();
    w.key("threads");
    w.value(rq.threads > 0 ? rq.threads : 1);
    w.key("results");
    w.begin_array();
}
inline JsonReply done(Writer& w) {
    w.end_array();
    w.end_object();
    JsonReply r;
    r.ok = true;
    r.text = w.take();
    return r;
}

template <class R>
void bookkeeping(Writer& w, const R& r) {
    w.key("eig_info");
    w.value(r.eig_info);
    w.key("n_eig");
    w.value(r.n_eig);
    w.key("n_dense_solve");
    w.value(r.n_dense_solve);
    w.key("max_eig_residual");
    w.value(r.max_eig_residual);
    w.key("max_solve_backward_error");
    w.value(r.max_solve_backward_error);
    w.key("timing");
    w.begin_object(true);
    w.key("total");
    w.value(r.timing.total);
    w.key("eig");
    w.value(r.timing.eig);
    w.key("linear");
    w.value(r.timing.linear);
    w.key("product");
    w.value(r.timing.product);
    w.end_object();
}

inline JsonReply run_planar(const Request& rq, int indent) {
    const std::vector<Result> res = run_solver([&] {
        std::vector<Point> pts;
        for (const PointSpec& p : rq.points) pts.push_back(Point{p.wavelength, p.theta_deg});
        return sweep(rq.stack, pts, rq.M, rq.polarization, rq.options);
    }, "M");
    Writer w(indent);
    head(w, rq);
    for (std::size_t k = 0; k < res.size(); ++k) {
        const Result& r = res[k];
        w.begin_object();
        w.key("wavelength");
        w.value(rq.points[k].wavelength);
        w.key("theta_deg");
        w.value(rq.points[k].theta_deg);
        w.key("R");
        put_complexes(w, r.R);
        w.key("T");
        put_complexes(w, r.T);
        w.key("DEr");
        put_reals(w, r.DEr);
        w.key("DEt");
        put_reals(w, r.DEt);
        w.key("R_total");
        w.value(r.reflectance());
        w.key("T_total");
        w.value(r.transmittance());
        bookkeeping(w, r);
        w.end_object();
    }
    return done(w);
}

inline void put_jones(Writer& w, const std::vector<Jones>& v) {
    w.begin_array(true);
    for (const Jones& j : v) {
        w.begin_array(true);
        put_complex(w, j.ss);
        put_complex(w, j.sp);
        put_complex(w, j.ps);
        put_complex(w, j.pp);
        w.end_array();
    }
    w.end_array();
}
inline void put_channels(Writer& w, const std::vector<JonesEfficiency>& v) {
    w.begin_array(true);
    for (const JonesEfficiency& e : v) {
        w.begin_array(true);
        w.value(e.ss);
        w.value(e.sp);
        w.value(e.ps);
        w.value(e.pp);
        w.end_array();
    }
    w.end_array();
}

inline JsonReply run_conical(const Request& rq, int indent) {
    const std::vector<ConicalResult> res = run_solver([&] {
        std::vector<ConicalPoint> pts;
        for (const PointSpec& p : rq.points) pts.push_back(ConicalPoint{p.wavelength, p.theta_deg, p.phi_deg});
        return sweep_conical(rq.stack, pts, rq.M, rq.options);
    }, "M");
    Writer w(indent);
    head(w, rq);
    for (std::size_t k = 0; k < res.size(); ++k) {
        const ConicalResult& r = res[k];
        const PointSpec& p = rq.points[k];
        w.begin_object();
        w.key("wavelength");
        w.value(p.wavelength);
        w.key("theta_deg");
        w.value(p.theta_deg);
        w.key("phi_deg");
        w.value(p.phi_deg);
        w.key("kx");
        put_reals(w, r.kx);
        w.key("ky");
        w.value(r.ky);
        w.key("phi_m_deg");
        put_reals(w, r.phi_m_deg);
        w.key("kzI");
        put_complexes(w, r.kzI);
        w.key("kzII");
        put_complexes(w, r.kzII);
        w.key("flux_r");
        put_reals(w, r.flux_r);
        w.key("flux_t");
        put_reals(w, r.flux_t);
        w.key("Jr");
        put_jones(w, r.Jr);
        w.key("Jt");
        put_jones(w, r.Jt);
        w.key("DEr_channels");
        put_channels(w, r.DEr);
        w.key("DEt_channels");
        put_channels(w, r.DEt);
        w.key("flux_unitarity");
        const std::array<cplx, 4> u = r.flux_unitarity();
        w.begin_array(true);
        for (const cplx& z : u) put_complex(w, z);
        w.end_array();
        if (rq.input != InputKind::None) {
            const ConicalPolarized c =
                rq.input == InputKind::Angles ? r.for_angles(rq.psi_deg, rq.delta_deg) : r.for_input(rq.Es, rq.Ep);
            w.key("polarized");
            w.begin_object();
            w.key("jones_in");
            w.begin_array(true);
            put_complex(w, c.Es);
            put_complex(w, c.Ep);
            w.end_array();
            const std::pair<const char*, const std::vector<cplx>*> amps[] = {{"Rs", &c.Rs}, {"Rp", &c.Rp}, {"Ts", &c.Ts},
                                                                               {"Tp", &c.Tp}};
            for (const auto& a : amps) {
                w.key(a.first);
                put_complexes(w, *a.second);
            }
            const std::pair<const char*, const std::vector<double>*> effs[] = {
                {"DEr", &c.DEr}, {"DEt", &c.DEt}, {"DEr_s", &c.DEr_s}, {"DEr_p", &c.DEr_p}, {"DEt_s", &c.DEt_s},
                {"DEt_p", &c.DEt_p}};
            for (const auto& e : effs) {
                w.key(e.first);
                put_reals(w, *e.second);
            }
            w.key("R_total");
            w.value(c.reflectance());
            w.key("T_total");
            w.value(c.transmittance());
            w.end_object();
        }
        bookkeeping(w, r);
        w.key("n_schur");
        w.value(r.n_schur);
        w.key("min_abs_lambda");
        w.value(r.min_abs_lambda);
        w.key("degenerate_mode_warning");
        w.value(r.degenerate_mode_warning);
        w.end_object();
    }
    return done(w);
}

inline const char* status_name(Status2D s) {
    return s == Status2D::Ok ? "ok" : s == Status2D::RayleighAnomaly ? "rayleigh_anomaly" : "numerical_failure";
}

inline JsonReply run_periodic2d(const Request& rq, int indent) {
    const Jones2D jones = rq.input == InputKind::Angles ? jones_from_angles_2d(rq.psi_deg, rq.delta_deg) : Jones2D{rq.Es, rq.Ep};
    const std::size_t count = static_cast<std::size_t>(2 * rq.M + 1) * static_cast<std::size_t>(2 * rq.My + 1);
    std::vector<Result2D> res;
    for (const PointSpec& p : rq.points)
        res.push_back(run_solver([&] {
            return solve2d(rq.stack2d, p.wavelength, p.theta_deg, p.phi_deg, rq.M, rq.My, jones, rq.options2d);
        }, "My"));
    Writer w(indent);
    head(w, rq);
    for (std::size_t k = 0; k < res.size(); ++k) {
        const Result2D& r = res[k];
        const PointSpec& p = rq.points[k];
        const bool solved = r.status == Status2D::Ok;
        const auto reals = [&](const char* key, const std::vector<double>& v) {
            w.key(key);
            if (solved) {
                put_reals(w, v);
                return;
            }
            w.begin_array(true);
            for (std::size_t j = 0; j < count; ++j) w.null();
            w.end_array();
        };
        const auto complexes = [&](const char* key, const std::vector<cplx>& v) {
            w.key(key);
            if (solved) {
                put_complexes(w, v);
                return;
            }
            w.begin_array(true);
            for (std::size_t j = 0; j < count; ++j) {
                w.begin_array(true);
                w.null();
                w.null();
                w.end_array();
            }
            w.end_array();
        };
        const auto flags = [&](const char* key, const std::vector<char>& v) {
            w.key(key);
            w.begin_array(true);
            for (char f : v) w.value(f != 0 ? 1 : 0);
            w.end_array();
        };
        w.begin_object();
        w.key("wavelength");
        w.value(p.wavelength);
        w.key("theta_deg");
        w.value(p.theta_deg);
        w.key("phi_deg");
        w.value(p.phi_deg);
        w.key("status");
        w.value(status_name(r.status));
        w.key("message");
        w.value(r.message);
        reals("Kx", r.Kx);
        reals("Ky", r.Ky);
        complexes("kzI", r.kzI);
        complexes("kzII", r.kzII);
        complexes("Rs", r.Rs);
        complexes("Rp", r.Rp);
        complexes("Ts", r.Ts);
        complexes("Tp", r.Tp);
        complexes("Ex_r", r.Ex_r);
        complexes("Ey_r", r.Ey_r);
        complexes("Ez_r", r.Ez_r);
        complexes("Ex_t", r.Ex_t);
        complexes("Ey_t", r.Ey_t);
        complexes("Ez_t", r.Ez_t);
        reals("DEr", r.DEr);
        reals("DEt", r.DEt);
        flags("prop_r", r.prop_r);
        flags("prop_t", r.prop_t);
        w.key("eig_info");
        w.value(r.eig_info);
        w.key("n_eig");
        w.value(r.n_eig);
        w.key("n_dense_solve");
        w.value(r.n_dense_solve);
        w.key("max_eig_residual");
        w.value(r.max_eig_residual);
        w.key("max_solve_backward_error");
        w.value(0.0);
        w.key("timing");
        w.begin_object(true);
        w.key("total");
        w.value(r.timing.total);
        w.key("operators");
        w.value(r.timing.operators);
        w.key("eig");
        w.value(r.timing.eig);
        w.key("linear");
        w.value(r.timing.linear);
        w.key("product");
        w.value(r.timing.product);
        w.end_object();
        w.end_object();
    }
    return done(w);
}

inline JsonReply error_reply(const std::string& message, const std::string& where, int indent) {
    Writer w(indent);
    w.begin_object();
    w.key("ok");
    w.value(false);
    w.key("error");
    w.value(message);
    w.key("where");
    w.value(where);
    w.end_object();
    JsonReply r;
    r.text = w.take();
    return r;
}

inline JsonReply answer(const std::string& text, int indent) {
    Request rq;
    {
        Value root;
        try {
            root = srcwa_amd::json::parse(text);
        } catch (const srcwa_amd::json::ParseError& e) {
            fail("", "the request is not valid JSON (line " + std::to_string(e.line()) + ", column " +
                         std::to_string(e.column()) + ", byte " + std::to_string(e.offset()) + "): " + e.reason());
        }
        rq = read_request(root);
    }
    if (rq.periodic2d()) return run_periodic2d(rq, indent);
    if (rq.conical()) return run_conical(rq, indent);
    return run_planar(rq, indent);
}

}  // namespace json_detail

// Request in, reply out, never throws.  indent < 0: the reply on one line; indent >= 0: one member per
// line, indented by indent spaces per level.
inline JsonReply solve_json_reply(const std::string& request, int indent = -1) {
    namespace jd = json_detail;
    try {
        try {
            return jd::answer(request, indent);
        } catch (const jd::RequestError& e) {
            return jd::error_reply(e.message, e.where, indent);
        } catch (const std::bad_alloc&) {
            return jd::error_reply("out of memory", "", indent);
        } catch (const std::exception& e) {
            return jd::error_reply(std::string("solver error: ") + e.what(), "", indent);
        } catch (...) {
            return jd::error_reply("unknown solver error", "", indent);
        }
    } catch (...) {
    }
    JsonReply fallback;
    try {
        fallback.text = "{\"ok\": false, \"error\": \"out of memory\", \"where\": \"\"}";
    } catch (...) {
    }
    return fallback;
}

inline std::string solve_json(const std::string& request) { return solve_json_reply(request).text; }

}  // namespace rcwa
}  // namespace srcwa_amd
