// Batched dense kernels in the shapes of the reference's hooks (custom_eig / custom_solve /
// custom_multiply, rcwa1d.hpp:122-136): one workgroup per item, items of any size in one launch.
#include "batch.hpp"
#include "eig.hpp"

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
#endif


// A is overwritten by the eigenvectors; scratch: Z, X (n x n each), ort, tmp (n each).
template <int QC>
__global__ __launch_bounds__(NT) void k_zgeev_batched(const DenseItem* __restrict__ items, cd* arena, int* info,
                                                       int max_iter) {
    const DenseItem it = items[blockIdx.x];
    const int n = it.n;
    const int64_t nn = (int64_t)n * n;
    cd* sc = arena + it.scratch;
    EigScratch s{sc + 2 * nn, sc + 2 * nn + n};
    const int r = wg_eig<QC>(n, arena + it.a, arena + it.w, sc, sc + nn, s, max_iter, kEps);
    if (threadIdx.x == 0) info[blockIdx.x] = r;
}

SRCWA_QC_DISPATCH(dense_qc, QcLaunch)
#ifdef SRCWA_QC_ONLY
void SRCWA_QC_FN(dense_qc)(const QcLaunch& a) {
    k_zgeev_batched<SRCWA_QC_ONLY><<<a.grid, NT, 0, a.stream>>>(a.items, a.ws, a.info, a.max_iter);
}
#endif

#if SRCWA_MAIN_TU
void launch_zgemm_batched(hipStream_t s, int count, const DenseItem* items, cd* arena) {
    k_zgemm_batched<<<count, NT, 0, s>>>(items, arena);
}
void launch_zgesv_batched(hipStream_t s, int count, const DenseItem* items, cd* arena, int* info) {
    k_zgesv_batched<<<count, NT, 0, s>>>(items, arena, info);
}
void launch_zgeev_batched(hipStream_t s, int count, const DenseItem* items, cd* arena, int* info, int max_iter,
                          int max_n) {
    QcLaunch a{};
    a.grid = count;
    a.stream = s;
    a.items = items;
    a.ws = arena;
    a.info = info;
    a.max_iter = max_iter;
    dense_qc(eig_columns_per_thread(max_n), a);
}
#endif
