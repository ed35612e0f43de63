// oracle/ref_decode.hpp -- test infrastructure, never part of the product: the flat stack records of
// include/srcwa.h decoded into the reference's own rcwa::Stack / rcwa::Stack2D (for the oracle wrappers
// built by oracle/Makefile).  Include after the reference's rcwa1d.hpp and rcwa2d.hpp.
#pragma once

#include <complex>
#include <cstddef>
#include <utility>
#include <vector>

namespace oracle_decode {

namespace rcwa = stablercwa::rcwa;
using stablercwa::cplx;

inline rcwa::Stack decode1d(const double* p) {
    rcwa::Stack s;
    s.period = p[0];
    s.nI = cplx{p[1], p[2]};
    s.nII = cplx{p[3], p[4]};
    const int L = static_cast<int>(p[5]);
    std::size_t k = 6;
    for (int j = 0; j < L; ++j) {
        const double d = p[k++];
        const int kind = static_cast<int>(p[k++]);
        if (kind == 0) {
            const cplx n{p[k], p[k + 1]};
            k += 2;
            s.layers.push_back(rcwa::Layer::uniform(d, n));
        } else {
            const int K = static_cast<int>(p[k++]);
            std::vector<double> e(p + k, p + k + K + 1);
            k += K + 1;
            std::vector<cplx> n;
            for (int i = 0; i < K; ++i) n.emplace_back(p[k + 2 * i], p[k + 2 * i + 1]);
            k += 2 * K;
            const int method = static_cast<int>(p[k++]);
            const int samples = static_cast<int>(p[k++]);
            s.layers.push_back(rcwa::Layer::grating(d, e, n,
                method ? rcwa::FourierMethod::FFT : rcwa::FourierMethod::Analytic, samples));
        }
    }
    return s;
}

inline rcwa::Stack2D decode2d(const double* p) {
    rcwa::Stack2D s;
    s.Lx = p[0];
    s.Ly = p[1];
    s.nI = p[2];
    s.nII = cplx{p[3], p[4]};
    const int L = static_cast<int>(p[5]);
    std::size_t k = 6;
    for (int j = 0; j < L; ++j) {
        const double d = p[k++];
        const int kind = static_cast<int>(p[k++]);
        if (kind == 0) {
            s.layers.push_back(rcwa::Layer2D::uniform(d, cplx{p[k], p[k + 1]}));
            k += 2;
        } else {
            const int Ky = static_cast<int>(p[k++]);
            std::vector<double> ye(p + k, p + k + Ky + 1);
            k += Ky + 1;
            std::vector<rcwa::XProfile2D> rows;
            for (int y = 0; y < Ky; ++y) {
                rcwa::XProfile2D prof;
                const int Kx = static_cast<int>(p[k++]);
                prof.edges.assign(p + k, p + k + Kx + 1);
                k += Kx + 1;
                for (int i = 0; i < Kx; ++i) prof.n_values.emplace_back(p[k + 2 * i], p[k + 2 * i + 1]);
                k += 2 * Kx;
                rows.push_back(std::move(prof));
            }
            s.layers.push_back(rcwa::Layer2D::patterned(d, ye, rows));
        }
    }
    return s;
}

}  // namespace oracle_decode
