// Batched dense kernels in the shapes of the reference's hooks (custom_eig / custom_solve /
// custom_multiply, rcwa1d.hpp:122-136): one workgroup per item, items of any size in one launch.
// Also the no-allocation eigensolve entry point (srcwa_zgeev_batched_noalloc) and its scratch array.
#include "batch.hpp"
#include "eig.hpp"
#if SRCWA_MIXED_PRECISION
#include "wg_la32.hpp"
#endif

#if SRCWA_MAIN_TU
#include <mutex>

#include "context.hpp"
#endif

#if SRCWA_MAIN_TU
__global__ __launch_bounds__(NT) void k_zgemm_batched(const DenseItem* __restrict__ items, cd* arena) {
    const DenseItem it = items[blockIdx.x];
    wg_gemm<GEMM_SET, false>(it.n, it.nrhs, it.k, arena + it.a, it.n, arena + it.b, it.k, arena + it.c, it.n);
}
#endif

#if SRCWA_MAIN_TU
// A is overwritten by its LU factors, B by the solution; scratch holds the pivots.
__global__ __launch_bounds__(NT) void k_zgesv_batched(const DenseItem* __restrict__ items, cd* arena, int* info) {
    const DenseItem it = items[blockIdx.x];
    int* piv = reinterpret_cast<int*>(arena + it.scratch);
    const int sing = wg_lu(it.n, arena + it.a, it.n, piv);
    if (!sing) wg_lu_solve(it.n, arena + it.a, it.n, piv, arena + it.b, it.n, it.nrhs);
    if (threadIdx.x == 0) info[blockIdx.x] = sing;
}

#if SRCWA_MIXED_PRECISION
// The mixed-precision form: A is preserved, B is overwritten by the solution; scratch holds the FP64
// fallback's factors (n x n) and pivots, then wg_refine_solve's scratch.  refine: 3 doubles per item
// (fell back to FP64, FP32 correction steps, final refined residual).
__global__ __launch_bounds__(NT) void k_zgesv_mixed_batched(const DenseItem* __restrict__ items, cd* arena, int* info,
                                                             double* refine) {
    const DenseItem it = items[blockIdx.x];
    const int64_t nn = (int64_t)it.n * it.n;
    cd* lu64 = arena + it.scratch;
    int* piv = reinterpret_cast<int*>(lu64 + nn);
    cd* scr = lu64 + nn + (it.n + 15) / 16 * 16;
    const RefineStats st = wg_refine_solve(it.n, arena + it.a, arena + it.b, it.nrhs, scr, lu64, piv);
    if (threadIdx.x == 0) {
        info[blockIdx.x] = st.singular;
        refine[3 * blockIdx.x] = st.fallback;
        refine[3 * blockIdx.x + 1] = st.iters;
        refine[3 * blockIdx.x + 2] = st.resid;
    }
}
#endif
#endif

// Item i: the matrix at abase + a (overwritten by the eigenvectors), eigenvalues at wbase + w, scratch
// (Z, X: n x n each; ort, tmp: n each) at sbase + scratch.
template <int QC>
__global__ __launch_bounds__(NT) void k_zgeev_batched(const DenseItem* __restrict__ items, cd* abase, cd* wbase,
                                                       cd* sbase, int* info, int max_iter) {
    const DenseItem it = items[blockIdx.x];
    const int n = it.n;
    const int64_t nn = (int64_t)n * n;
    cd* sc = sbase + it.scratch;
    EigScratch s{sc + 2 * nn, sc + 2 * nn + n};
    const int r = wg_eig<QC>(n, abase + it.a, wbase + it.w, sc, sc + nn, s, max_iter, kEps);
    if (threadIdx.x == 0) info[blockIdx.x] = r;
}

SRCWA_QC_DISPATCH(dense_qc, QcLaunch)
#ifdef SRCWA_QC_ONLY
void SRCWA_QC_FN(dense_qc)(const QcLaunch& a) {
    k_zgeev_batched<SRCWA_QC_ONLY><<<a.grid, NT, 0, a.stream>>>(a.items, a.ws, a.wbase, a.sbase, a.info, a.max_iter);
}
#endif

#if SRCWA_MAIN_TU
void launch_zgemm_batched(hipStream_t s, int count, const DenseItem* items, cd* arena) {
    k_zgemm_batched<<<count, NT, 0, s>>>(items, arena);
}
void launch_zgesv_batched(hipStream_t s, int count, const DenseItem* items, cd* arena, int* info) {
    k_zgesv_batched<<<count, NT, 0, s>>>(items, arena, info);
}
#if SRCWA_MIXED_PRECISION
void launch_zgesv_mixed_batched(hipStream_t s, int count, const DenseItem* items, cd* arena, int* info, double* refine) {
    k_zgesv_mixed_batched<<<count, NT, 0, s>>>(items, arena, info, refine);
}
#endif
void launch_zgeev_batched(hipStream_t s, int count, const DenseItem* items, cd* arena, int* info, int max_iter,
                          int max_n) {
    QcLaunch a{};
    a.grid = count;
    a.stream = s;
    a.items = items;
    a.ws = arena;
    a.wbase = arena;
    a.sbase = arena;
    a.info = info;
    a.max_iter = max_iter;
    dense_qc(eig_columns_per_thread(max_n), a);
}

// ------------------------------------------------------------------------------------------------
// The no-allocation entry point.  Its scratch (the item descriptors and each item's Z, X, ort, tmp)
// lives in this file-scope device array; every GPU the library runs on holds its own copy until the
// process ends.  Host code never uses the array's name as an address: the device address of each GPU's
// copy comes from hipGetSymbolAddress, once per device, when a context is created on it.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_lock;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int64_t round16(int64_t x) { return (x + 15) & ~int64_t(15); }
int64_t round256(int64_t x) { return (x + 255) & ~int64_t(255); }
}  // namespace

int dense_scratch_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return static_cast<int>(hipErrorInvalidDevice);
    std::lock_guard<std::mutex> g(g_scratch_lock);
    if (g_scratch_ptr[device]) return static_cast<int>(hipSuccess);
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[device] = p;
    return static_cast<int>(e);
}

// Descriptors of the packed items from their orders (one workgroup; an exclusive scan per 256 items):
// matrix i at sum_{k<i} n_k^2, eigenvalues at sum_{k<i} n_k, scratch at sum_{k<i} round16(2 n_k^2 + 2 n_k).
__global__ __launch_bounds__(NT) void k_eig_items(int count, const int* __restrict__ n, DenseItem* __restrict__ items) {
    __shared__ int64_t s_a[NT], s_w[NT], s_s[NT];
    const int t = threadIdx.x;
    int64_t base_a = 0, base_w = 0, base_s = 0;
    for (int c0 = 0; c0 < count; c0 += NT) {
        const int i = c0 + t;
        const int64_t ni = i < count ? n[i] : 0;
        const int64_t va = ni * ni, vw = ni, vs = i < count ? ((2 * ni * ni + 2 * ni + 15) & ~int64_t(15)) : 0;
        s_a[t] = va;
        s_w[t] = vw;
        s_s[t] = vs;
        __syncthreads();
        for (int off = 1; off < NT; off <<= 1) {
            const int64_t xa = t >= off ? s_a[t - off] : 0, xw = t >= off ? s_w[t - off] : 0,
                          xs = t >= off ? s_s[t - off] : 0;
            __syncthreads();
            s_a[t] += xa;
            s_w[t] += xw;
            s_s[t] += xs;
            __syncthreads();
        }
        if (i < count) {
            DenseItem it;
            it.n = static_cast<int>(ni);
            it.nrhs = 0;
            it.k = 0;
            it.a = base_a + s_a[t] - va;
            it.b = 0;
            it.c = 0;
            it.w = base_w + s_w[t] - vw;
            it.scratch = base_s + s_s[t] - vs;
            items[i] = it;
        }
        base_a += s_a[NT - 1];
        base_w += s_w[NT - 1];
        base_s += s_s[NT - 1];
        __syncthreads();
    }
}

extern "C" int srcwa_zgeev_batched_noalloc(srcwa_context* ctx, int count, const int* n, const double* A, double* w,
                                           double* V, int* info) {
    if (!ctx || count < 1 || !n || !A || !w || !V || !info) {
        srcwa_set_last_error("srcwa_zgeev_batched_noalloc: empty or null inputs");
        return SRCWA_EINVAL;
    }
    int64_t total_a = 0, total_s = 0;
    int max_n = 0;
    for (int i = 0; i < count; ++i) {
        if (n[i] < 1 || n[i] > EIG_MAX_N_HOST) {
            srcwa_set_last_error("srcwa_zgeev_batched_noalloc: order out of range");
            return SRCWA_EINVAL;
        }
        const int64_t ni = n[i];
        total_a += ni * ni;
        total_s += round16(2 * ni * ni + 2 * ni);
        max_n = n[i] > max_n ? n[i] : max_n;
    }
    const int64_t n_bytes = round256(int64_t(count) * int64_t(sizeof(int)));
    const int64_t item_bytes = round256(int64_t(count) * int64_t(sizeof(DenseItem)));
    const int64_t need = n_bytes + item_bytes + total_s * int64_t(sizeof(cd));
    if (need > int64_t(NOALLOC_SCRATCH_BYTES)) {
        srcwa_set_last_error("srcwa_zgeev_batched_noalloc: the batch needs " + std::to_string(need) +
                             " bytes of scratch, above the " + std::to_string(NOALLOC_SCRATCH_BYTES >> 20) +
                             " MiB this entry point holds per device; split the batch");
        return SRCWA_ENOMEM;
    }
    auto fail = [](hipError_t e, const char* what) {
        srcwa_set_last_error(std::string("srcwa_zgeev_batched_noalloc: ") + what + ": " + hipGetErrorString(e));
        return SRCWA_EDEVICE;
    };
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail(e, "hipSetDevice");
    char* base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_scratch_lock);
        base = static_cast<char*>(g_scratch_ptr[ctx->device]);
    }
    if (!base) {
        srcwa_set_last_error("srcwa_zgeev_batched_noalloc: no scratch on this device (context not created)");
        return SRCWA_EDEVICE;
    }
    int* d_n = reinterpret_cast<int*>(base);
    DenseItem* d_items = reinterpret_cast<DenseItem*>(base + n_bytes);
    cd* d_scratch = reinterpret_cast<cd*>(base + n_bytes + item_bytes);
    hipStream_t s = ctx->stream;
    e = hipMemcpyAsync(d_n, n, count * sizeof(int), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return fail(e, "hipMemcpyAsync (orders)");
    if (V != A) {
        e = hipMemcpyAsync(V, A, total_a * sizeof(cd), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return fail(e, "hipMemcpyAsync (A to V)");
    }
    k_eig_items<<<1, NT, 0, s>>>(count, d_n, d_items);
    QcLaunch a{};
    a.grid = count;
    a.stream = s;
    a.items = d_items;
    a.ws = reinterpret_cast<cd*>(V);
    a.wbase = reinterpret_cast<cd*>(w);
    a.sbase = d_scratch;
    a.info = info;
    a.max_iter = 30;
    dense_qc(eig_columns_per_thread(max_n), a);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(e, "kernel launch");
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(e, "hipStreamSynchronize");
    return SRCWA_OK;
}
#endif
