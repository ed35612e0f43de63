// Prints the reference's ten conical Jones-matrix tables (tests/conical_reference.inc of the reference)
// as JSON, each with its stack in the flat record encoding.  Built and run only by
// make_conical_fixtures.py in a container that holds the reference tree.
#include <cstdio>
#include <limits>
#include <stdexcept>

#include "stablercwa/rcwa1d.hpp"
#include "conical_reference.inc"

namespace rcwa = stablercwa::rcwa;

static void print_stack(const rcwa::Stack& s) {
    std::printf("[%.17g, %.17g, %.17g, %.17g, %.17g, %d", s.period, s.nI.real(), s.nI.imag(), s.nII.real(),
                s.nII.imag(), static_cast<int>(s.layers.size()));
    for (const rcwa::Layer& l : s.layers) {
        std::printf(", %.17g", l.d);
        if (l.homogeneous) {
            std::printf(", 0, %.17g, %.17g", l.n.real(), l.n.imag());
        } else {
            std::printf(", 1, %d", static_cast<int>(l.n_values.size()));
            for (double e : l.edges) std::printf(", %.17g", e);
            for (const auto& v : l.n_values) std::printf(", %.17g, %.17g", v.real(), v.imag());
            std::printf(", 0, 0");
        }
    }
    std::printf("]");
}

int main() {
    std::printf("{\"source\": \"reference tests/conical_reference.inc, tolerance 1e-9 max(1,|J|)\", \"cases\": [\n");
    bool first = true;
    for (const ConicalRefCase& c : kConicalRefCases) {
        if (!first) std::printf(",\n");
        first = false;
        std::printf("{\"id\": %d, \"name\": \"%s\", \"wavelength\": %.17g, \"theta\": %.17g, \"phi\": %.17g, \"M\": %d, \"stack\": ",
                    c.id, c.name, c.wavelength, c.theta_deg, c.phi_deg, c.M);
        print_stack(conical_ref_stack(c.id));
        std::printf(", \"Jr\": [");
        for (int i = 0; i < c.count; ++i)
            for (int k = 0; k < 8; ++k) std::printf("%s%.17g", (i || k) ? ", " : "", c.orders[i].jr[k]);
        std::printf("], \"Jt\": [");
        for (int i = 0; i < c.count; ++i)
            for (int k = 0; k < 8; ++k) std::printf("%s%.17g", (i || k) ? ", " : "", c.orders[i].jt[k]);
        std::printf("]}");
    }
    std::printf("\n]}\n");
    return 0;
}
