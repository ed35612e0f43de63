// Drives the C++ host API (include/stablercwa_amd/rcwa.hpp) for the GPU tests: reads one problem per
// line on stdin, builds the C++ Stack / Stack2D from the flat record, calls rcwa::solve, sweep,
// solve_conical or solve2d, and prints the results with 17 significant digits.
//
//   planar  <pol> <M> <wavelength> <theta> <n> <flat record (n doubles)>
//   sweep   <pol> <M> <points> (<wavelength> <theta>)* <n> <flat record>
//   planar_diag <pol> <M> <wavelength> <theta> <n> <flat record>   (Options::diagnostics = true)
//   conical <M> <wavelength> <theta> <phi> <n> <flat record>
//   crossed <M1> <M2> <wavelength> <theta> <phi> <Es.re> <Es.im> <Ep.re> <Ep.im> <laurent> <n> <flat record>
//   fields  <pol> <M> <wavelength> <theta> <nz> <z>* <nx> <y> <n> <flat record>   (solve_with_fields, field_at,
//           field_map; prints the 6 harmonic components order-major per depth, then the 6 grid components)
//   conical_fields <M> <wavelength> <theta> <phi> <Es.re> <Es.im> <Ep.re> <Ep.im> <nz> <z>* <nx> <y> <n> <record>
//           (solve_conical_with_fields, field_at and field_map for that Jones vector)
//
// Output per problem: "ok" and labelled arrays, or "invalid_argument <message>" / "unsupported <message>".
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "stablercwa_amd/fields.hpp"
#include "stablercwa_amd/rcwa.hpp"

namespace rcwa = srcwa_amd::rcwa;
using srcwa_amd::cplx;

static std::vector<double> read_record(std::istream& in) {
    int n = 0;
    in >> n;
    std::vector<double> v(n);
    for (double& x : v) in >> x;
    return v;
}

static rcwa::Stack planar_stack(const std::vector<double>& f) {
    if (f.size() < 6) throw std::invalid_argument("driver: malformed stack record");
    rcwa::Stack s;
    s.period = f[0];
    s.nI = cplx{f[1], f[2]};
    s.nII = cplx{f[3], f[4]};
    const int L = static_cast<int>(f[5]);
    std::size_t k = 6;
    for (int j = 0; j < L; ++j) {
        const double d = f[k];
        const int kind = static_cast<int>(f[k + 1]);
        k += 2;
        if (kind == 0) {
            s.layers.push_back(rcwa::Layer::uniform(d, cplx{f[k], f[k + 1]}));
            k += 2;
        } else {
            const int K = static_cast<int>(f[k]);
            std::vector<double> edges(f.begin() + k + 1, f.begin() + k + 2 + K);
            std::vector<cplx> n;
            for (int i = 0; i < K; ++i) n.emplace_back(f[k + 2 + K + 2 * i], f[k + 3 + K + 2 * i]);
            k += 1 + (K + 1) + 2 * K;
            const auto method = static_cast<rcwa::FourierMethod>(static_cast<int>(f[k]));
            const int samples = static_cast<int>(f[k + 1]);
            k += 2;
            s.layers.push_back(rcwa::Layer::grating(d, edges, n, method, samples));
        }
    }
    return s;
}

static rcwa::Stack2D crossed_stack(const std::vector<double>& f) {
    if (f.size() < 6) throw std::invalid_argument("driver: malformed stack record");
    rcwa::Stack2D s;
    s.Lx = f[0];
    s.Ly = f[1];
    s.nI = f[2];
    s.nII = cplx{f[3], f[4]};
    const int L = static_cast<int>(f[5]);
    std::size_t k = 6;
    for (int j = 0; j < L; ++j) {
        const double d = f[k];
        const int kind = static_cast<int>(f[k + 1]);
        k += 2;
        if (kind == 0) {
            s.layers.push_back(rcwa::Layer2D::uniform(d, cplx{f[k], f[k + 1]}));
            k += 2;
            continue;
        }
        const int Ky = static_cast<int>(f[k]);
        std::vector<double> y_edges(f.begin() + k + 1, f.begin() + k + 2 + Ky);
        k += 1 + (Ky + 1);
        std::vector<rcwa::XProfile2D> rows;
        for (int y = 0; y < Ky; ++y) {
            const int Kx = static_cast<int>(f[k]);
            rcwa::XProfile2D p;
            p.edges.assign(f.begin() + k + 1, f.begin() + k + 2 + Kx);
            for (int i = 0; i < Kx; ++i) p.n_values.emplace_back(f[k + 2 + Kx + 2 * i], f[k + 3 + Kx + 2 * i]);
            k += 1 + (Kx + 1) + 2 * Kx;
            rows.push_back(std::move(p));
        }
        s.layers.push_back(rcwa::Layer2D::patterned(d, y_edges, rows));
    }
    return s;
}

static void put(const char* label, const std::vector<double>& v) {
    std::printf("%s", label);
    for (double x : v) std::printf(" %.17g", x);
    std::printf("\n");
}
static void put(const char* label, const std::vector<cplx>& v) {
    std::printf("%s", label);
    for (const cplx& x : v) std::printf(" %.17g %.17g", x.real(), x.imag());
    std::printf("\n");
}

static void print_planar(const rcwa::Result& r) {
    std::printf("ok %d %zu\n", r.eig_info, r.DEr.size());
    put("DEr", r.DEr);
    put("DEt", r.DEt);
    put("R", r.R);
    put("T", r.T);
}

int main() {
    std::string line;
    while (std::getline(std::cin, line)) {
        if (line.empty()) continue;
        std::istringstream in(line);
        std::string kind;
        in >> kind;
        try {
            if (kind == "planar") {
                int pol, M;
                double wl, th;
                in >> pol >> M >> wl >> th;
                const auto st = planar_stack(read_record(in));
                print_planar(rcwa::solve(st, wl, th, M, pol ? rcwa::Polarization::TM : rcwa::Polarization::TE));
            } else if (kind == "planar_diag") {
                int pol, M;
                double wl, th;
                in >> pol >> M >> wl >> th;
                const auto st = planar_stack(read_record(in));
                rcwa::Options opt;
                opt.diagnostics = true;
                const rcwa::Result r = rcwa::solve(st, wl, th, M, pol ? rcwa::Polarization::TM : rcwa::Polarization::TE, opt);
                print_planar(r);
                std::printf("diag %.17g %.17g %d %d %.17g %.17g %.17g %.17g\n", r.max_eig_residual,
                            r.max_solve_backward_error, r.n_eig, r.n_dense_solve, r.timing.total, r.timing.eig,
                            r.timing.linear, r.timing.product);
            } else if (kind == "sweep") {
                int pol, M, P;
                in >> pol >> M >> P;
                std::vector<rcwa::Point> pts(P);
                for (auto& p : pts) in >> p.wavelength >> p.theta_deg;
                const auto st = planar_stack(read_record(in));
                const auto rs = rcwa::sweep(st, pts, M, pol ? rcwa::Polarization::TM : rcwa::Polarization::TE);
                std::printf("sweep %zu\n", rs.size());
                for (const auto& r : rs) print_planar(r);
            } else if (kind == "conical") {
                int M;
                double wl, th, ph;
                in >> M >> wl >> th >> ph;
                const auto st = planar_stack(read_record(in));
                const rcwa::ConicalResult r = rcwa::solve_conical(st, wl, th, ph, M);
                std::printf("ok %d %zu\n", r.eig_info, r.Jr.size());
                std::vector<cplx> jr, jt;
                std::vector<double> er, et;
                for (std::size_t j = 0; j < r.Jr.size(); ++j) {
                    jr.insert(jr.end(), {r.Jr[j].ss, r.Jr[j].sp, r.Jr[j].ps, r.Jr[j].pp});
                    jt.insert(jt.end(), {r.Jt[j].ss, r.Jt[j].sp, r.Jt[j].ps, r.Jt[j].pp});
                    er.insert(er.end(), {r.DEr[j].ss, r.DEr[j].sp, r.DEr[j].ps, r.DEr[j].pp});
                    et.insert(et.end(), {r.DEt[j].ss, r.DEt[j].sp, r.DEt[j].ps, r.DEt[j].pp});
                }
                put("Jr", jr);
                put("Jt", jt);
                put("DEr", er);
                put("DEt", et);
                put("flux_r", r.flux_r);
                put("flux_t", r.flux_t);
                const auto u = r.flux_unitarity();
                put("unitarity", std::vector<cplx>(u.begin(), u.end()));
                const auto s = r.for_angles(90.0);
                put("s_DEr", s.DEr);
            } else if (kind == "crossed") {
                int M1, M2, laurent;
                double wl, th, ph, a, b, c, d;
                in >> M1 >> M2 >> wl >> th >> ph >> a >> b >> c >> d >> laurent;
                const auto st = crossed_stack(read_record(in));
                rcwa::Options2D opt;
                opt.factorization = laurent ? rcwa::Factorization2D::Laurent : rcwa::Factorization2D::Li;
                const rcwa::Result2D r = rcwa::solve2d(st, wl, th, ph, M1, M2, {cplx{a, b}, cplx{c, d}}, opt);
                std::printf("ok %d %d %d\n", r.eig_info, static_cast<int>(r.status), r.N);
                put("DEr", r.DEr);
                put("DEt", r.DEt);
                put("Rs", r.Rs);
                put("Rp", r.Rp);
                put("Ts", r.Ts);
                put("Tp", r.Tp);
                put("Kx", r.Kx);
                put("Ky", r.Ky);
            } else if (kind == "fields") {
                int pol, M, nz, nx;
                double wl, th, y;
                in >> pol >> M >> wl >> th >> nz;
                srcwa_amd::RealVector zs(nz);
                for (double& z : zs) in >> z;
                in >> nx >> y;
                const auto st = planar_stack(read_record(in));
                const auto res = rcwa::solve_with_fields(st, wl, th, M, pol ? rcwa::Polarization::TM : rcwa::Polarization::TE);
                std::printf("ok %d %zu %d\n", res.result.eig_info, res.result.DEr.size(), res.fields.available() ? 1 : 0);
                put("DEr", res.result.DEr);
                if (res.fields.available()) {
                    const rcwa::PlanarFieldContext copy = res.fields;       // copies share the capture
                    const auto h = rcwa::field_at(copy, zs);
                    const auto g = rcwa::field_map(res, zs, nx, y);
                    std::vector<cplx> hv, gv;
                    for (int c = 0; c < 6; ++c) {
                        hv.insert(hv.end(), h.components[c].data(), h.components[c].data() + h.components[c].rows() * h.components[c].cols());
                        gv.insert(gv.end(), g.components[c].data(), g.components[c].data() + g.components[c].rows() * g.components[c].cols());
                    }
                    put("harm", hv);
                    put("grid", gv);
                    put("kzI", h.metadata.kzI);
                    std::printf("z_bottom %.17g\n", h.metadata.z_bottom);
                } else {
                    try {
                        (void)rcwa::field_at(res.fields, zs);
                        std::printf("field_at did not throw\n");
                    } catch (const std::logic_error& e) {
                        std::printf("logic_error %s\n", e.what());
                    }
                }
            } else if (kind == "conical_fields") {
                int M, nz, nx;
                double wl, th, ph, a, b, c, d, y;
                in >> M >> wl >> th >> ph >> a >> b >> c >> d >> nz;
                srcwa_amd::RealVector zs(nz);
                for (double& z : zs) in >> z;
                in >> nx >> y;
                const auto st = planar_stack(read_record(in));
                const auto res = rcwa::solve_conical_with_fields(st, wl, th, ph, M);
                std::printf("ok %d %zu %d\n", res.result.eig_info, res.result.kx.size(), res.fields.available() ? 1 : 0);
                if (res.fields.available()) {
                    const rcwa::JonesVector jones{{cplx{a, b}, cplx{c, d}}};
                    const auto h = rcwa::field_at(res, zs, jones);
                    const auto g = rcwa::field_map(res.fields, zs, nx, y, jones);
                    std::vector<cplx> hv, gv;
                    for (int k = 0; k < 6; ++k) {
                        hv.insert(hv.end(), h.components[k].data(), h.components[k].data() + h.components[k].rows() * h.components[k].cols());
                        gv.insert(gv.end(), g.components[k].data(), g.components[k].data() + g.components[k].rows() * g.components[k].cols());
                    }
                    put("harm", hv);
                    put("grid", gv);
                    std::printf("ky %.17g\n", h.metadata.ky);
                    std::printf("schur_layers");
                    for (const auto& l : h.metadata.layers)
                        if (l.has_modes) std::printf(" %d", l.propagation == rcwa::FieldPropagation::Schur ? 1 : 0);
                    std::printf("\n");
                }
            } else {
                std::printf("invalid_argument unknown problem kind\n");
            }
        } catch (const srcwa_amd::Unsupported& e) {
            std::printf("unsupported %s\n", e.what());
        } catch (const std::invalid_argument& e) {
            std::printf("invalid_argument %s\n", e.what());
        } catch (const std::exception& e) {
            std::printf("error %s\n", e.what());
        }
        std::printf("end\n");
        std::fflush(stdout);
    }
    return 0;
}
