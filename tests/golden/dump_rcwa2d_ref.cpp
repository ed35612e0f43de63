// Prints the reference's crossed-grating reference cases (tests/rcwa2d_reference.inc of the reference)
// as JSON: inputs with the stack in the flat record encoding, per-order far-field arrays and, where the
// table holds them, the layer operators.  Built and run only by make_rcwa2d_fixtures.py in a container
// that holds the reference tree.
#include <cstdio>
#include <limits>
#include <stdexcept>

#include "stablercwa/rcwa2d_types.hpp"
#include "rcwa2d_reference.inc"

namespace rcwa = stablercwa::rcwa;

static void print_stack(const rcwa::Stack2D& s) {
    std::printf("[%.17g, %.17g, %.17g, %.17g, %.17g, %d", s.Lx, s.Ly, s.nI, s.nII.real(), s.nII.imag(),
                static_cast<int>(s.layers.size()));
    for (const rcwa::Layer2D& l : s.layers) {
        std::printf(", %.17g", l.d);
        if (l.homogeneous) {
            std::printf(", 0, %.17g, %.17g", l.n.real(), l.n.imag());
        } else {
            std::printf(", 1, %d", static_cast<int>(l.x_profiles.size()));
            for (double y : l.y_edges) std::printf(", %.17g", y);
            for (const rcwa::XProfile2D& p : l.x_profiles) {
                std::printf(", %d", static_cast<int>(p.n_values.size()));
                for (double x : p.edges) std::printf(", %.17g", x);
                for (const auto& v : p.n_values) std::printf(", %.17g, %.17g", v.real(), v.imag());
            }
        }
    }
    std::printf("]");
}

static void arr(const char* name, const Rcwa2dRefCase& c, double (*get)(const Rcwa2dRefOrder&)) {
    std::printf(", \"%s\": [", name);
    for (int i = 0; i < c.count; ++i) std::printf("%s%.17g", i ? ", " : "", get(c.orders[i]));
    std::printf("]");
}
static void carr(const char* name, const Rcwa2dRefCase& c, const double* (*get)(const Rcwa2dRefOrder&)) {
    std::printf(", \"%s\": [", name);
    for (int i = 0; i < c.count; ++i) {
        const double* v = get(c.orders[i]);
        std::printf("%s%.17g, %.17g", i ? ", " : "", v[0], v[1]);
    }
    std::printf("]");
}

int main() {
    std::printf("{\"source\": \"reference tests/rcwa2d_reference.inc (frozen Python solver), |a-b| <= atol + rtol |b|\", \"cases\": [\n");
    for (int k = 0; k < kRcwa2dRefCaseCount; ++k) {
        const Rcwa2dRefCase& c = kRcwa2dRefCases[k];
        if (k) std::printf(",\n");
        std::printf("{\"id\": %d, \"name\": \"%s\", \"wavelength\": %.17g, \"theta\": %.17g, \"phi\": %.17g, \"M1\": %d, "
                    "\"M2\": %d, \"jones\": [%.17g, %.17g, %.17g, %.17g], \"laurent\": %d, \"atol\": %.17g, \"rtol\": %.17g, "
                    "\"status\": %d, \"stack\": ",
                    c.id, c.name, c.wavelength, c.theta_deg, c.phi_deg, c.M1, c.M2, c.jones[0], c.jones[1], c.jones[2],
                    c.jones[3], c.laurent, c.far_atol, c.far_rtol, c.status);
        print_stack(rcwa2d_ref_stack(c.id));
        std::printf(", \"failure_orders\": [");
        for (int i = 0; i < c.failure_count; ++i)
            std::printf("%s[%d, %d]", i ? ", " : "", c.failure_orders[i][0], c.failure_orders[i][1]);
        std::printf("]");
        if (c.status == 0) {
            arr("Kx", c, [](const Rcwa2dRefOrder& o) { return o.Kx; });
            arr("Ky", c, [](const Rcwa2dRefOrder& o) { return o.Ky; });
            arr("DEr", c, [](const Rcwa2dRefOrder& o) { return o.DEr; });
            arr("DEt", c, [](const Rcwa2dRefOrder& o) { return o.DEt; });
            carr("Rs", c, [](const Rcwa2dRefOrder& o) { return static_cast<const double*>(o.Rs); });
            carr("Rp", c, [](const Rcwa2dRefOrder& o) { return static_cast<const double*>(o.Rp); });
            carr("Ts", c, [](const Rcwa2dRefOrder& o) { return static_cast<const double*>(o.Ts); });
            carr("Tp", c, [](const Rcwa2dRefOrder& o) { return static_cast<const double*>(o.Tp); });
            carr("Ex_r", c, [](const Rcwa2dRefOrder& o) { return static_cast<const double*>(o.Ex_r); });
            carr("Ey_r", c, [](const Rcwa2dRefOrder& o) { return static_cast<const double*>(o.Ey_r); });
            carr("Ex_t", c, [](const Rcwa2dRefOrder& o) { return static_cast<const double*>(o.Ex_t); });
            carr("Ey_t", c, [](const Rcwa2dRefOrder& o) { return static_cast<const double*>(o.Ey_t); });
        }
        std::printf(", \"operators\": [");
        for (int i = 0; i < c.operator_count; ++i) {
            const Rcwa2dRefOperators& op = c.operators[i];
            std::printf("%s{\"layer\": %d, \"N\": %d", i ? ", " : "", op.layer, op.N);
            const double* m[3] = {op.eps, op.cx, op.cy};
            const char* nm[3] = {"eps", "cx", "cy"};
            for (int q = 0; q < 3; ++q) {
                std::printf(", \"%s\": [", nm[q]);
                for (int e = 0; e < 2 * op.N * op.N; ++e) std::printf("%s%.17g", e ? ", " : "", m[q][e]);
                std::printf("]");
            }
            std::printf("}");
        }
        std::printf("]}");
    }
    std::printf("\n]}\n");
    return 0;
}
