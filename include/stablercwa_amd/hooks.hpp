// GPU implementations of the reference's operator hooks, header-only, over the batched C-ABI kernels of
// srcwa.h (link with -lsrcwa_amd).  For callers that keep the reference's own solver and want the GPU behind
// its plug-in points: Options::custom_eig, custom_multiply, custom_solve (rcwa1d.hpp:122-136; the same
// shapes in Options2D, rcwa2d_types.hpp:89-95).
//
//   #include "stablercwa/rcwa1d.hpp"             // the reference's headers first
//   #include "stablercwa_amd/hooks.hpp"
//   stablercwa::rcwa::Options opt;
//   srcwa_amd::hooks::use_gpu(opt);               // all three hooks, as mkl::use_mkl (mkl_backend.hpp:386-391)
//   opt.custom_eig = srcwa_amd::hooks::gpu_eig<stablercwa::Matrix, stablercwa::EigResult>();   // or one
//
// Matrix is any type with rows(), cols(), data() (interleaved complex, column-major) and Matrix(rows, cols),
// as the reference's stablercwa::Matrix (matrix.hpp:34-75); EigResult has values, vectors and info
// (eig_types.hpp:33-40).  The hook contracts:
//   gpu_eig       eigenvalues and unit 2-norm right eigenvectors (order and phase free); info > 0 (an
//                 eigenvalue not converged within max_iter sweeps) leaves vectors empty
//   gpu_multiply  A.rows() x B.cols()
//   gpu_solve     A.cols() x B.cols(); a singular A (exact zero pivot) throws std::runtime_error, which the
//                 reference's solver turns into a numerical failure (NaN results, rcwa1d.hpp:272-275)
//   all three     wrong operand shapes throw std::invalid_argument; a HIP runtime error throws
//                 std::runtime_error
// Calls are serialised per device, so the hooks may be called from the reference's pool threads
// (parallel_layers, sweep).  Each call moves one problem across PCIe and cannot batch across layers or
// sweep points: the plan API (srcwa_plan*_create, or include/stablercwa_amd/rcwa.hpp) is the fast path.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>

#include "../srcwa.h"

namespace srcwa_amd {
namespace hooks {
namespace detail {

struct Device {
    srcwa_context* ctx = nullptr;
    std::mutex lock;     // one call at a time on this device's stream
};

// The hooks' context on a device, created on first use and kept until the process ends.
inline Device& device(int d) {
    static std::mutex m;
    static std::map<int, std::unique_ptr<Device>> devices;
    std::lock_guard<std::mutex> g(m);
    std::unique_ptr<Device>& slot = devices[d];
    if (!slot) {
        auto dev = std::make_unique<Device>();
        if (srcwa_context_create(d, &dev->ctx) != SRCWA_OK) throw std::runtime_error(srcwa_last_error());
        slot = std::move(dev);
    }
    return *slot;
}

inline void check(int code) {
    if (code == SRCWA_OK) return;
    if (code == SRCWA_EINVAL) throw std::invalid_argument(srcwa_last_error());
    throw std::runtime_error(srcwa_last_error());
}

template <class P>
const double* in(const P* p) {
    return reinterpret_cast<const double*>(p);
}
template <class P>
double* out(P* p) {
    return reinterpret_cast<double*>(p);
}

template <class F>
struct eig_signature;
template <class R, class M>
struct eig_signature<std::function<R(const M&)>> {
    using Result = R;
    using Mat = M;
};

}  // namespace detail

// custom_eig: eig_qr's contract (EigResult, eig_types.hpp:33-40) computed by the batched GPU eigensolver
// (srcwa_zgeev_batched_iter); max_iter is EigOptions::max_iter.
template <class Matrix, class EigResult>
std::function<EigResult(const Matrix&)> gpu_eig(int device = 0, int max_iter = 30) {
    return [device, max_iter](const Matrix& a) -> EigResult {
        if (a.rows() != a.cols()) throw std::invalid_argument("gpu_eig: matrix must be square");
        const int n = a.rows();
        EigResult r;
        r.values.resize(static_cast<std::size_t>(n));
        if (n == 0) return r;
        Matrix v(n, n);
        int info = 0;
        detail::Device& d = detail::device(device);
        {
            std::lock_guard<std::mutex> g(d.lock);
            detail::check(srcwa_zgeev_batched_iter(d.ctx, 1, &n, detail::in(a.data()), detail::out(r.values.data()),
                                                   detail::out(v.data()), &info, max_iter));
        }
        r.info = info;
        if (info == 0) r.vectors = std::move(v);     // no vectors when an eigenvalue did not converge
        return r;
    };
}

// custom_multiply: C = A B (srcwa_zgemm_batched).
template <class Matrix>
std::function<Matrix(const Matrix&, const Matrix&)> gpu_multiply(int device = 0) {
    return [device](const Matrix& a, const Matrix& b) -> Matrix {
        if (a.cols() != b.rows()) throw std::invalid_argument("gpu_multiply: inner dimensions differ");
        const int m = a.rows(), n = b.cols(), k = a.cols();
        Matrix c(m, n);
        if (m == 0 || n == 0 || k == 0) return c;
        detail::Device& d = detail::device(device);
        std::lock_guard<std::mutex> g(d.lock);
        detail::check(srcwa_zgemm_batched(d.ctx, 1, &m, &n, &k, detail::in(a.data()), detail::in(b.data()),
                                          detail::out(c.data())));
        return c;
    };
}

// custom_solve: X = A^{-1} B by partial-pivot LU (srcwa_zgesv_batched, the reference's lu.hpp:267-443
// algorithm); an exact zero pivot throws std::runtime_error.
template <class Matrix>
std::function<Matrix(const Matrix&, const Matrix&)> gpu_solve(int device = 0) {
    return [device](const Matrix& a, const Matrix& b) -> Matrix {
        if (a.rows() != a.cols() || a.rows() != b.rows()) throw std::invalid_argument("gpu_solve: shape mismatch");
        const int n = a.rows(), nrhs = b.cols();
        Matrix x(n, nrhs);
        if (n == 0 || nrhs == 0) return x;
        int info = 0;
        detail::Device& d = detail::device(device);
        {
            std::lock_guard<std::mutex> g(d.lock);
            detail::check(srcwa_zgesv_batched(d.ctx, 1, &n, &nrhs, detail::in(a.data()), detail::in(b.data()),
                                              detail::out(x.data()), &info));
        }
        if (info != 0) throw std::runtime_error("gpu_solve: singular matrix (exact zero pivot)");
        return x;
    };
}

// All three hooks of an Options / Options2D (types deduced from its custom_eig member).
template <class Options>
void use_gpu(Options& options, int device = 0) {
    using S = detail::eig_signature<decltype(options.custom_eig)>;
    options.custom_eig = gpu_eig<typename S::Mat, typename S::Result>(device);
    options.custom_multiply = gpu_multiply<typename S::Mat>(device);
    options.custom_solve = gpu_solve<typename S::Mat>(device);
}

}  // namespace hooks
}  // namespace srcwa_amd
