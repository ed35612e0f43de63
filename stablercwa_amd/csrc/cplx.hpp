// Complex FP64 arithmetic for gfx950 device code.
//
// Matrices are column-major arrays of `cd` (interleaved re, im), element (i, j) at p[i + j*ld],
// the same layout as the reference's Matrix (include/stablercwa/matrix.hpp:1-5).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

struct __attribute__((aligned(16))) cd {
    double x, y;
};

#define DEV __device__ __forceinline__

// Member-wise load / store of one complex value.  An aggregate copy (`a = *p`) is a 16-byte memcpy to
// the compiler, which can keep register arrays of cd from being promoted out of private memory; these
// are two double accesses that the backend merges into one 16-byte access.
DEV cd ldc(const cd* p) { return cd{p->x, p->y}; }
DEV void stc(cd* p, cd a) {
    p->x = a.x;
    p->y = a.y;
}

DEV cd mk(double x, double y = 0.0) { return cd{x, y}; }
DEV cd operator+(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
DEV cd operator-(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
DEV cd operator-(cd a) { return cd{-a.x, -a.y}; }
DEV cd operator*(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
DEV cd operator*(double s, cd a) { return cd{s * a.x, s * a.y}; }
DEV cd operator*(cd a, double s) { return cd{s * a.x, s * a.y}; }
DEV cd conj(cd a) { return cd{a.x, -a.y}; }
DEV bool is_zero(cd a) { return a.x == 0.0 && a.y == 0.0; }
DEV bool finite(cd a) { return isfinite(a.x) && isfinite(a.y); }
// |Re| + |Im|, the EISPACK magnitude (reference matrix.hpp:32).
DEV double abs1(cd a) { return fabs(a.x) + fabs(a.y); }
DEV double cabs(cd a) { return hypot(a.x, a.y); }
DEV double norm2(cd a) { return a.x * a.x + a.y * a.y; }

// a / b by Smith's method (scaled, no intermediate overflow for moderate operands).
DEV cd cdiv(cd a, cd b) {
    if (fabs(b.x) >= fabs(b.y)) {
        if (b.x == 0.0 && b.y == 0.0) return cd{a.x / 0.0, a.y / 0.0};
        const double r = b.y / b.x, den = b.x + b.y * r;
        return cd{(a.x + a.y * r) / den, (a.y - a.x * r) / den};
    }
    const double r = b.x / b.y, den = b.x * r + b.y;
    return cd{(a.x * r + a.y) / den, (a.y * r - a.x) / den};
}
DEV cd cdiv(cd a, double s) { return cd{a.x / s, a.y / s}; }

// Principal square root: Re >= 0, Im carries the sign of Im z (also for -0.0).
DEV cd csqrt(cd z) {
    const double x = z.x, y = z.y;
    if (x == 0.0 && y == 0.0) return cd{0.0, y};
    if (isinf(y)) return cd{INFINITY, y};
    const double t = sqrt(0.5 * (fabs(x) + hypot(x, y)));
    if (x >= 0.0) return cd{t, y / (2.0 * t)};
    return cd{fabs(y) / (2.0 * t), copysign(t, y)};
}

DEV cd cexp(cd z) {
    const double e = exp(z.x);
    double s, c;
    sincos(z.y, &s, &c);
    return cd{e * c, e * s};
}

// m = f * 2^e with f in [0.5, 1); 0 for zero or non-finite m (reference eig_qr.hpp "binary_exponent").
DEV int binary_exponent(double m) {
    int e = 0;
    if (m > 0.0 && isfinite(m)) frexp(m, &e);
    return e;
}
DEV cd scale2(cd z, int e) { return cd{ldexp(z.x, e), ldexp(z.y, e)}; }

// z / |z| and |z|, computed after a power-of-two rescale (reference eig_qr.hpp unit_phase).
DEV cd unit_phase(cd z, double& modulus) {
    const int e = binary_exponent(fmax(fabs(z.x), fabs(z.y)));
    z = scale2(z, -e);
    const double r = hypot(z.x, z.y);
    modulus = ldexp(r, e);
    return cd{z.x / r, z.y / r};
}

constexpr double kEps = 2.220446049250313e-16;
constexpr double kPi = 3.141592653589793238462643383279502884;
