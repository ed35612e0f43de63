// Host runtime of libsrcwa_amd.so: contexts, launch plans, HBM-footprint admission into chunks,
// input validation with the reference's error conventions, and the C-ABI of include/srcwa.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <map>
#include <complex>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/srcwa.h"
#include "batch.hpp"
#include <numeric>

#include "context.hpp"
#include "tuned.hpp"
#include "cplx.hpp"

// kernel launchers (rcwa1d.hip, dense.hip)
void launch_planar(hipStream_t s, int n_inst, int n_tasks, const Inst1D* insts, const LayerTask* tasks,
                   const double* stacks, const int64_t* layouts, cd* ws, int* task_info, Out1D out, int max_iter, int min_n, int max_n,
                   int part, int flags, hipEvent_t* ev, const OpsArgs& ops);
void launch_conical(hipStream_t s, int n_inst, int n_tasks, const Inst1D* insts, const LayerTask* tasks,
                    const double* stacks, const int64_t* layouts, cd* ws, int* task_info, Out1D out, int max_iter, int min_n, int max_n,
                    int part, int flags, hipEvent_t* ev, const int* eig_order, const OpsArgs& ops);
void launch_crossed(hipStream_t s, int n_inst, int n_tasks, const Inst2D* insts, const LayerTask* tasks,
                    const double* stacks, const int64_t* layouts, cd* ws, int* task_info, Out1D out, int max_iter, int min_n, int max_n,
                    int part, int flags, hipEvent_t* ev, const int* eig_order, const OpsArgs& ops);
void launch_fields1d(hipStream_t s, int n_groups, const Inst1D* inst, const double* stacks, const int64_t* layouts,
                     const cd* ws, const cd* fields, const cd* R, const cd* T, const FieldGroup1D* groups,
                     const int* cols, const double* zs, int nz, cd* out, cd* scratch);
void launch_fieldmap1d(hipStream_t s, const Inst1D* inst, const double* stacks, const int64_t* layouts, const cd* ws,
                       const cd* H, int nz, int nx, double kyy, cd* out);
void launch_fields_conical(hipStream_t s, int n_groups, const Inst1D* inst, const double* stacks,
                           const int64_t* layouts, const cd* ws, const cd* fields, const FieldGroup1D* groups,
                           const int* cols, const double* zs, int nz, cd j0, cd j1, cd* out, cd* scratch);
void launch_zgemm_batched(hipStream_t s, int count, const DenseItem* items, cd* arena);
void launch_zgesv_batched(hipStream_t s, int count, const DenseItem* items, cd* arena, int* info);
#if SRCWA_MIXED_PRECISION
void launch_zgesv_mixed_batched(hipStream_t s, int count, const DenseItem* items, cd* arena, int* info, double* refine);
#endif
constexpr const char* kMixedOff =
    "mixed-precision star products are compiled out of this build (SRCWA_MIXED_PRECISION 0): their first GPU run "
    "ended in a memory fault whose cause is not found yet";
void launch_zgeev_batched(hipStream_t s, int count, const DenseItem* items, cd* arena, int* info, int max_iter,
                          int max_n);

namespace {

thread_local std::string g_error;

struct Failure {
    int code;
    std::string msg;
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) throw Failure{SRCWA_ENOMEM, std::string(what) + ": out of device memory"};
        throw Failure{SRCWA_EDEVICE, std::string(what) + ": " + hipGetErrorString(e)};
    }
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return SRCWA_OK;
    } catch (const Failure& e) {
        g_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_error = "host allocation failed";
        return SRCWA_ENOMEM;
    } catch (const std::exception& e) {
        g_error = e.what();
        return SRCWA_EDEVICE;
    }
}

void invalid(const std::string& m) { throw Failure{SRCWA_EINVAL, m}; }

}  // namespace

void srcwa_set_last_error(const std::string& msg) { g_error = msg; }

static_assert(int(SRCWA_PLAN_DIAGNOSTICS) == int(PLAN_DIAGNOSTICS), "plan flag of the C-ABI and the kernels");
static_assert(int(SRCWA_PLAN_MIXED) == int(PLAN_MIXED), "plan flag of the C-ABI and the kernels");
static_assert(int(SRCWA_PLAN_POINT_OPS) == int(PLAN_POINT_OPS), "plan flag of the C-ABI and the kernels");

// A device arena carved first-fit into 256-byte aligned blocks.  Plans created on a context that has
// a pool take every device buffer from it, so many plans can come and go (a request server) without
// hipMalloc / hipFree, and a plan that does not fit fails with SRCWA_ENOMEM instead of growing.
struct srcwa_pool {
    int device = 0;
    char* base = nullptr;
    size_t bytes = 0;
    size_t used = 0;
    std::mutex lock;
    std::map<size_t, size_t> free_blocks;    // offset -> size, coalesced

    static size_t round(size_t n) { return (n + 255) & ~size_t(255); }
    void* get(size_t n) {
        n = round(n);
        std::lock_guard<std::mutex> g(lock);
        for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
            if (it->second < n) continue;
            const size_t off = it->first, left = it->second - n;
            free_blocks.erase(it);
            if (left) free_blocks[off + n] = left;
            used += n;
            return base + off;
        }
        return nullptr;
    }
    void put(void* p, size_t n) {
        n = round(n);
        std::lock_guard<std::mutex> g(lock);
        size_t off = static_cast<size_t>(static_cast<char*>(p) - base);
        used -= n;
        auto next = free_blocks.lower_bound(off);
        if (next != free_blocks.end() && off + n == next->first) {
            n += next->second;
            next = free_blocks.erase(next);
        }
        if (next != free_blocks.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += n;
                return;
            }
        }
        free_blocks[off] = n;
    }
};

namespace {

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    srcwa_pool* pool = nullptr;     // allocate from this arena instead of hipMalloc
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            if (pool) pool->put(p, n * sizeof(T));
            else (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        if (pool) {
            p = static_cast<T*>(pool->get(count * sizeof(T)));
            if (!p) throw Failure{SRCWA_ENOMEM, "device pool exhausted"};
        } else {
            hip_check(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)), "hipMalloc");
        }
        n = count;
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        alloc(v.size());
        if (!v.empty())
            hip_check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpy H2D");
    }
    size_t bytes() const { return n * sizeof(T); }
};

int64_t align16(int64_t x) { return (x + 15) & ~int64_t(15); }

// Per chunk, the chunk-local indices of its layer tasks ordered by decreasing eigen order (stable), so the
// eigen kernels start the longest eigensolves first instead of leaving them for the end of the launch.
template <class Chunk, class Inst>
std::vector<int> longest_first(const std::vector<Chunk>& chunks, const std::vector<LayerTask>& tasks,
                               const std::vector<Inst>& insts) {
    std::vector<int> order(std::max<size_t>(1, tasks.size()), 0);
    for (const Chunk& ch : chunks) {
        auto first = order.begin() + ch.t0, last = order.begin() + ch.t1;
        for (int t = ch.t0; t < ch.t1; ++t) order[t] = t - ch.t0;
        std::stable_sort(first, last, [&](int a, int b) {
            return insts[ch.i0 + tasks[ch.t0 + a].inst].N > insts[ch.i0 + tasks[ch.t0 + b].inst].N;
        });
    }
    return order;
}

// pool bytes of a plan's diagnostics buffers (per-instance records, per-task records, residuals)
int64_t diag_bytes(int n_inst, int n_tasks) {
    const auto b = [](size_t x) { return static_cast<int64_t>((x + 255) & ~size_t(255)); };
    const size_t t = static_cast<size_t>(std::max(1, n_tasks));
    return b(size_t(DIAG_WORDS) * n_inst * sizeof(double)) + b(t * sizeof(TaskDiag)) + b(2 * t * sizeof(double));
}

// ---- validation (rcwa1d.hpp:188-229) ------------------------------------------------------------
bool finite_c(double re, double im) { return std::isfinite(re) && std::isfinite(im); }

void check_index(double re, double im) {
    if (!finite_c(re, im) || std::hypot(re, im) == 0.0 || re < 0.0 || im < 0.0)
        invalid("rcwa: refractive index must be finite, nonzero, with Re n >= 0 and Im n >= 0");
}

struct LayerInfo {
    bool grating;
    bool uniform_values;   // a grating whose segment indices are all exactly equal
    double d;
    int method = 0;        // gratings: 0 analytic, 1 FFT
    int samples = 0;       // gratings: fft_samples (0 automatic)
    int64_t rec = 0;       // the layer's words after its thickness (kind onward) in the stack record
    int64_t rec_words = 0;
};

constexpr int kMaxFFTSamples = 4194304;

// resolve_fft_samples (fourier.hpp:26-44): the sample count of an FFT grating at truncation order M, with
// the reference's messages for the settings it rejects.
int resolve_fft_samples(int M, int samples) {
    if (M > (kMaxFFTSamples - 1) / 4) invalid("rcwa: FFT truncation requires more than 4194304 samples");
    if (samples == 0) {
        const long long target = std::max(8192LL, 64LL * (2LL * M + 1));
        if (target > kMaxFFTSamples) invalid("rcwa: automatic FFT exceeds 4194304 samples");
    }
    const int n = fft_samples_for(M, samples);
    if (n <= 4 * M) invalid("rcwa: fft_samples must be at least 4*M+1");
    return n;
}

// cd words of FFT sample scratch each grating layer of an instance at order M needs (0 when analytic)
int64_t fft_scratch_words(const LayerInfo& li, int M) {
    return li.grating && li.method == 1 ? 2 * int64_t(resolve_fft_samples(M, li.samples)) : 0;
}

// Parses and validates one planar stack record; returns the per-layer kinds.
std::vector<LayerInfo> parse_stack1d(const double* p, int64_t avail, int M) {
    if (avail < 6) invalid("rcwa: malformed stack record");
    const double period = p[0];
    const int L = static_cast<int>(p[5]);
    if (!(p[5] >= 0.0) || L > 128) invalid("rcwa: a stack holds 0..128 layers");
    std::vector<LayerInfo> out;
    int64_t k = 6;
    auto need = [&](int64_t n) {
        if (k + n > avail) invalid("rcwa: malformed stack record");
    };
    for (int j = 0; j < L; ++j) {
        need(2);
        const double d = p[k];
        const int kind = static_cast<int>(p[k + 1]);
        const int64_t start = k + 1;
        k += 2;
        if (kind == 0) {
            need(2);
            if (!std::isfinite(d) || d < 0.0) invalid("rcwa: layer thickness must be finite and >= 0");
            check_index(p[k], p[k + 1]);
            k += 2;
            out.push_back({false, true, d});
        } else if (kind == 1) {
            need(1);
            const int K = static_cast<int>(p[k]);
            if (K < 1 || K > 100000) invalid("rcwa: a grating layer needs K indices and K+1 edges");
            need(1 + (K + 1) + 2 * K + 2);
            const double* e = p + k + 1;
            const double* v = e + K + 1;
            // validate_fourier_settings (fourier.hpp:17-23)
            const double method_word = v[2 * K], samples_word = v[2 * K + 1];
            if (method_word != 0.0 && method_word != 1.0) invalid("rcwa: invalid Fourier method");
            if (!(samples_word >= 0.0 && samples_word <= kMaxFFTSamples) || samples_word != std::floor(samples_word))
                invalid("rcwa: fft_samples must be 0..4194304 (0 for automatic)");
            const int method = static_cast<int>(method_word), samples = static_cast<int>(samples_word);
            if (!std::isfinite(d) || d < 0.0) invalid("rcwa: layer thickness must be finite and >= 0");
            for (int i = 0; i <= K; ++i)
                if (!std::isfinite(e[i]) || (i > 0 && e[i] <= e[i - 1]))
                    invalid("rcwa: grating edges must be finite and strictly increasing");
            if (std::abs((e[K] - e[0]) - period) > 1e-10 * period)
                invalid("rcwa: grating edges must span exactly one period");
            bool same = true;
            for (int i = 0; i < K; ++i) {
                check_index(v[2 * i], v[2 * i + 1]);
                same = same && v[2 * i] == v[0] && v[2 * i + 1] == v[1];
            }
            k += 1 + (K + 1) + 2 * K + 2;
            out.push_back({true, same, d, method, samples, start, k - start});
        } else {
            invalid("rcwa: unknown layer kind");
        }
    }
    (void)M;
    return out;
}

void check_setup(double period, double wavelength, double nIr, double nIi, double theta_deg, int M, int pol) {
    if (M < 0 || M > (std::numeric_limits<int>::max() - 1) / 4) invalid("rcwa: invalid truncation order M");
    if (pol != 0 && pol != 1) invalid("rcwa: polarization must be TE or TM");
    if (!std::isfinite(wavelength) || wavelength <= 0.0 || !std::isfinite(period) || period <= 0.0)
        invalid("rcwa: wavelength and period must be finite and positive");
    if (!std::isfinite(theta_deg) || std::abs(theta_deg) >= 90.0)
        invalid("rcwa: incidence angle must lie strictly between -90 and 90 degrees");
    check_index(nIr, nIi);
    if (nIi != 0.0 || nIr <= 0.0) invalid("rcwa: the incident medium must have a real positive index");
}

// Shared operator sets: layer tasks whose layer operators cannot depend on the point (the same layer
// words after the thickness, the same period or cell and truncation, the same polarization) are grouped
// by those values, bit for bit, and every group of two or more tasks gets one set, computed once per
// execution; single tasks keep computing their own.
struct OpSharing {
    std::map<std::vector<uint64_t>, int> index;
    std::vector<int> task_key;         // per task of the plan, in task order
    std::vector<int> count;
    std::vector<LayerOps> rep;         // a plan-wide (instance, layer) of each key
    std::vector<int64_t> words;
    bool enabled = true;

    static void push(std::vector<uint64_t>& key, double v) {
        uint64_t u;
        std::memcpy(&u, &v, sizeof u);
        key.push_back(u);
    }
    // key: the values the operators depend on; words: the set's size in cd words
    void add(const std::vector<uint64_t>& key, int inst, int layer, int64_t set_words) {
        auto it = index.find(key);
        if (it == index.end()) {
            it = index.emplace(key, static_cast<int>(count.size())).first;
            count.push_back(0);
            rep.push_back(LayerOps{inst, layer, 0});
            words.push_back(set_words);
        }
        ++count[it->second];
        task_key.push_back(it->second);
    }
    // Gives each shared group its set (tasks in the order add() saw them); returns the operator buffer's
    // size in cd words.
    int64_t assign(std::vector<LayerTask>& tasks, std::vector<LayerOps>& descs) const {
        descs.clear();
        if (!enabled) return 0;
        std::vector<int> op(count.size(), -1);
        int64_t off = 0;
        for (size_t k = 0; k < count.size(); ++k)
            if (count[k] >= 2) {
                op[k] = static_cast<int>(descs.size());
                descs.push_back(LayerOps{rep[k].inst, rep[k].layer, off});
                off += (words[k] + 15) & ~int64_t(15);
            }
        for (size_t t = 0; t < tasks.size(); ++t) tasks[t].op = op[task_key[t]];
        return off;
    }
};

// device bytes of a plan's shared operator sets (as the pool rounds them)
int64_t ops_bytes(const std::vector<LayerOps>& descs, int64_t words) {
    if (descs.empty()) return 0;
    return static_cast<int64_t>(srcwa_pool::round(descs.size() * sizeof(LayerOps)) +
                                srcwa_pool::round(static_cast<size_t>(words) * sizeof(cd)) +
                                srcwa_pool::round(descs.size() * sizeof(int)));
}

struct ParsedStacks {
    std::vector<std::vector<LayerInfo>> infos;
    std::vector<double> words;
};

// Validates every flat stack record and copies the records' extent.
ParsedStacks parse_all(int n_stacks, const double* stacks, const int64_t* stack_offsets) {
    ParsedStacks out;
    std::vector<int64_t> sorted(stack_offsets, stack_offsets + n_stacks);
    for (int64_t o : sorted)
        if (o < 0) invalid("rcwa: malformed stack offsets");
    std::sort(sorted.begin(), sorted.end());
    auto record_end = [&](int64_t off) {
        auto it = std::upper_bound(sorted.begin(), sorted.end(), off);
        return it == sorted.end() ? int64_t(-1) : *it;
    };
    out.infos.resize(n_stacks);
    int64_t stack_words = 0;
    for (int s = 0; s < n_stacks; ++s) {
        const int64_t off = stack_offsets[s];
        const int64_t end = record_end(off);
        const int64_t avail = end > 0 ? end - off : (int64_t(1) << 40);
        out.infos[s] = parse_stack1d(stacks + off, avail, 0);
        int64_t k = 6;
        const double* p = stacks + off;
        for (size_t j = 0; j < out.infos[s].size(); ++j) {
            const int kind = static_cast<int>(p[k + 1]);
            k += 2;
            if (kind == 0) k += 2;
            else {
                const int K = static_cast<int>(p[k]);
                k += 1 + (K + 1) + 2 * K + 2;
            }
        }
        stack_words = std::max(stack_words, off + k);
    }
    out.words.assign(stacks, stacks + stack_words);
    return out;
}

// ---- operator-kind simulation (mirrors run_planar_recursion) ------------------------------------
enum Kind { KZ = 0, KD = 1, KX = 2 };
int kmul(int a, int b) { return (a == KZ || b == KZ) ? KZ : (a == KD && b == KD) ? KD : KX; }
int kadd(int a, int b) { return b == KZ ? a : a == KZ ? b : (a == KD && b == KD) ? KD : KX; }

struct PlanarCost {
    int64_t bump_words;
    double flops_eig, flops_other;
};

// general[r]: region r (0 incidence, j+1 layer j, L+1 exit) needs an eigensolve.
PlanarCost planar_cost(int N, const std::vector<bool>& general, int pol, bool conical = false) {
    const int L = static_cast<int>(general.size()) - 2;
    const double n = N, n3 = n * n * n;
    const double LU = 8.0 / 3.0 * n3, SOLVE_N = 8.0 * n3, GEMM = 8.0 * n3;
    PlanarCost c{0, 0.0, 0.0};
    int64_t persist = 2 * align16(int64_t(N) * N);
    int r_worst = KZ, r = KZ;
    for (int b = 2 * L; b >= 1; --b) {
        const int j = b / 2;
        int tdd, rdu, rud, tuu, rdu_w, rud_w;
        if (b % 2 == 0) {
            const bool dense = general[j] || general[j + 1];
            tdd = rdu = rud = tuu = dense ? KX : KD;
            rdu_w = rdu;
            rud_w = rud;
            if (dense) {
                if (general[j + 1]) c.flops_other += 2 * LU + 2 * SOLVE_N;   // matching solves
                c.flops_other += LU + SOLVE_N + 3 * GEMM;                     // T+^-1, Rdu, Tdd, Rud
            }
        } else {
            tdd = tuu = KD;
            rdu = rud = KZ;
            rdu_w = rud_w = conical ? KZ : KD;   // only planar finite layers have the zero-root reflection
        }
        // flops along the expected kinds
        const int k = kmul(rdu, r);
        if (rdu == KX && r == KX) c.flops_other += GEMM;
        const int g = k == KZ ? tdd : k == KD ? kmul(KD, tdd) : KX;
        if (k == KX) c.flops_other += LU + SOLVE_N;
        if (r == KX && g == KX) c.flops_other += GEMM;
        const int rg = kmul(r, g);
        if (tuu == KX && rg == KX) c.flops_other += GEMM;
        r = kadd(rud, kmul(tuu, rg));
        // memory along the worst-case kinds
        const int kw = kmul(rdu_w, r_worst);
        const int gw = kw == KZ ? tdd : kw == KD ? kmul(KD, tdd) : KX;
        persist += align16(gw == KD ? N : int64_t(N) * N);
        r_worst = kadd(rud_w, kmul(tuu, kmul(r_worst, gw)));
    }
    // top block
    if (general[0] || general[1]) {
        if (general[1]) c.flops_other += 2 * LU + 2 * SOLVE_N;
        c.flops_other += LU + SOLVE_N + GEMM;
        if (r == KX) c.flops_other += GEMM + LU + 8.0 * n * n;
    }
    for (int j = 1; j <= L; ++j)
        if (general[j]) {
            if (conical) {
                // two N x N eigensolves (N = half the mode count), 2 LU inverses, B C, C W2, coupling
                const double h = n / 2, h3 = h * h * h;
                c.flops_eig += 2 * 100.0 * h3;
                c.flops_other += 2 * (8.0 / 3.0 * h3) + 2 * 8.0 * h3 + 3 * 8.0 * h3;
            } else {
                c.flops_eig += 100.0 * n3;
                if (pol == 1) c.flops_other += 2 * LU + 2 * SOLVE_N + GEMM;   // TM assembly + V
            }
        }
    const int64_t NN = int64_t(N) * N;
    c.bump_words = persist + 16 * NN + 64 * int64_t(N) + 4096;
    return c;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
struct PlanChunk {
    int i0, i1, t0, t1;
};

struct srcwa_plan {
    srcwa_context* ctx = nullptr;
    int kind = 1;
    int max_n = 1;                // largest harmonic count N of any instance (sizes the eigen kernels)
    int min_n = 1 << 30;          // smallest N (whether the small-matrix eigen kernel runs at all)
    int n_inst = 0, n_tasks = 0;
    int64_t total_orders = 0;
    std::vector<PlanChunk> chunks;
    DevBuf<Inst1D> insts;
    DevBuf<Inst2D> insts2d;
    DevBuf<LayerTask> tasks;
    DevBuf<double> stacks;
    DevBuf<int64_t> layouts;
    DevBuf<cd> ws;
    DevBuf<int> task_info;
    DevBuf<int> eig_order;          // conical / crossed: each chunk's tasks, largest eigensolve first
    DevBuf<LayerOps> op_descs;      // shared operator sets (LayerOps), their buffer and failure words
    DevBuf<cd> ops;
    DevBuf<int> ops_info;
    int n_ops = 0;
    DevBuf<double> R, T, DEr, DEt;
    DevBuf<double> geometry;      // conical / crossed: GEO_WORDS per order
    DevBuf<int> status, eig_info;
    DevBuf<double> diag;            // DIAG_WORDS per instance
    DevBuf<TaskDiag> task_diag;     // per layer task
    DevBuf<double> item_res;        // conical: two eigen residuals per layer task
    // planar plans with PLAN_FIELDS: the layer amplitudes a+, a- (2 L N complex per problem), a per-problem
    // flag that the capture is whole, and host copies of what the samplers need
    DevBuf<cd> fields;
    DevBuf<int> field_ok;
    std::vector<Inst1D> h_insts;
    std::vector<std::vector<LayerInfo>> h_layers;   // per problem
    std::vector<double> h_stack_words;
    bool executed = false;
    int flags = 0;                  // SRCWA_PLAN_* flags (from the context at creation)
    // 5 timing events per chunk: front start, after assembly, after eigensolves (front stream); back
    // start, back end (back stream); then the final event on the front stream
    std::vector<hipEvent_t> ev;
    std::vector<hipEvent_t> front_done, back_done;   // ordering only
    int64_t region_words = 0;        // workspace of one chunk; two regions alternate between chunks
    hipGraph_t graph = nullptr;      // captured launch sequence (srcwa_plan_set_graph)
    hipGraphExec_t graph_exec = nullptr;
    bool timed = false;
    int max_iter = 30;               // QR sweeps per eigenvalue (EigOptions::max_iter, eig_types.hpp)
    double flops[2] = {0.0, 0.0};
    ~srcwa_plan() {
        // uploads and launches still queued on the context's streams may touch the buffers below, which
        // return to a shared pool (another plan may get them at once) when the members are destroyed
        if (ctx) {
            (void)hipSetDevice(ctx->device);
            for (hipStream_t st : {ctx->stream, ctx->back})
                if (st) (void)hipStreamSynchronize(st);
        }
        for (std::vector<hipEvent_t>* v : {&ev, &front_done, &back_done})
            for (hipEvent_t& e : *v)
                if (e) (void)hipEventDestroy(e);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
    // every device buffer of the plan comes from the context's pool, when it has one
    void bind_pool(srcwa_pool* p) {
        insts.pool = p;
        insts2d.pool = p;
        tasks.pool = p;
        stacks.pool = p;
        layouts.pool = p;
        ws.pool = p;
        task_info.pool = p;
        eig_order.pool = p;
        op_descs.pool = p;
        ops.pool = p;
        ops_info.pool = p;
        R.pool = p;
        T.pool = p;
        DEr.pool = p;
        DEt.pool = p;
        geometry.pool = p;
        status.pool = p;
        eig_info.pool = p;
        diag.pool = p;
        task_diag.pool = p;
        item_res.pool = p;
        fields.pool = p;
        field_ok.pool = p;
    }
    void alloc_diag() {
        diag.alloc(size_t(DIAG_WORDS) * n_inst);
        task_diag.alloc(std::max(1, n_tasks));
        item_res.alloc(2 * size_t(std::max(1, n_tasks)));
    }
    void make_events() {
        const size_t c = chunks.size();
        ev.assign(5 * c + 1, nullptr);
        for (hipEvent_t& e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
        if (c > 1) {
            front_done.assign(c, nullptr);
            back_done.assign(c, nullptr);
            for (std::vector<hipEvent_t>* v : {&front_done, &back_done})
                for (hipEvent_t& e : *v) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
            if (!ctx->back) hip_check(hipStreamCreateWithFlags(&ctx->back, hipStreamNonBlocking), "hipStreamCreate");
        }
    }
    int64_t bytes() const {
        return insts.bytes() + insts2d.bytes() + tasks.bytes() + stacks.bytes() + layouts.bytes() + ws.bytes() + task_info.bytes() +
               R.bytes() + T.bytes() + DEr.bytes() + DEt.bytes() + status.bytes() + eig_info.bytes() + geometry.bytes() +
               diag.bytes() + task_diag.bytes() + item_res.bytes() + eig_order.bytes() + op_descs.bytes() + ops.bytes() +
               ops_info.bytes();
    }
    // uploads the shared operator sets of a plan (none: nothing allocated)
    void upload_ops(const std::vector<LayerOps>& descs, int64_t words, hipStream_t st) {
        n_ops = static_cast<int>(descs.size());
        if (n_ops == 0) return;
        op_descs.upload(descs, st);
        ops.alloc(static_cast<size_t>(words));
        ops_info.alloc(static_cast<size_t>(n_ops));
    }
};

namespace {
// Device-to-host copy on the plan's own stream (the legacy null stream would order it against every
// other stream of the process), followed by a wait for that stream only.
void copy_out(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, const char* what, srcwa_plan* plan) {
    hip_check(hipMemcpyAsync(dst, src, bytes, kind, plan->ctx->stream), what);
    hip_check(hipStreamSynchronize(plan->ctx->stream), what);
}

// ---- hook-shaped dense batches ------------------------------------------------------------------
template <class Body>
int dense_batch(srcwa_context* ctx, Body&& body) {
    return guarded([&] {
        if (!ctx) invalid("null context");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        body(ctx->stream);
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

}  // namespace

namespace {
// Enqueues every chunk of the plan on the context's stream(s); with timing, the stage events too.
void enqueue(srcwa_plan* plan, bool timing) {
    hipStream_t front = plan->ctx->stream;
    const size_t nc = plan->chunks.size();
    // One chunk: every stage on the context stream.  Several chunks: the eigensolves of chunk c+1
    // (front stream) overlap the S-matrix work of chunk c (back stream); chunks alternate between two
    // workspace regions, so chunk c+2 starts only after chunk c's back half has released its region.
    hipStream_t back = nc > 1 ? plan->ctx->back : front;
    auto launch = [&](size_t c, hipStream_t st, int part) {
        const PlanChunk& ch = plan->chunks[c];
        Out1D out{plan->R.p, plan->T.p, plan->DEr.p, plan->DEt.p, plan->status.p + ch.i0, plan->eig_info.p + ch.i0,
                  plan->kind == 1 ? nullptr : plan->geometry.p, plan->diag.p + int64_t(DIAG_WORDS) * ch.i0,
                  plan->task_diag.p + ch.t0, plan->item_res.p + 2 * int64_t(ch.t0)};
        if (plan->kind <= 2 && (plan->flags & PLAN_FIELDS)) {
            out.fields = reinterpret_cast<double*>(plan->fields.p);
            out.field_ok = plan->field_ok.p + ch.i0;
        }
        cd* ws = plan->ws.p + (nc > 1 ? static_cast<int64_t>(c % 2) * plan->region_words : 0);
        hipEvent_t* ev = timing ? plan->ev.data() + 5 * c : nullptr;
        const int ni = ch.i1 - ch.i0, nt = ch.t1 - ch.t0;
        // the shared operator sets are computed on the front stream ahead of the first chunk; every later
        // stage that reads them is ordered after it (same stream, or the back stream after front_done)
        OpsArgs ops;
        ops.descs = plan->op_descs.p;
        ops.ops = plan->ops.p;
        ops.info = plan->ops_info.p;
        ops.n = plan->n_ops;
        ops.insts = plan->kind == 3 ? static_cast<const void*>(plan->insts2d.p) : static_cast<const void*>(plan->insts.p);
        ops.compute = c == 0;
        if (plan->kind == 1)
            launch_planar(st, ni, nt, plan->insts.p + ch.i0, plan->tasks.p + ch.t0, plan->stacks.p, plan->layouts.p,
                          ws, plan->task_info.p + ch.t0, out, plan->max_iter, plan->min_n, plan->max_n, part, plan->flags, ev,
                          ops);
        else if (plan->kind == 2)
            launch_conical(st, ni, nt, plan->insts.p + ch.i0, plan->tasks.p + ch.t0, plan->stacks.p, plan->layouts.p,
                           ws, plan->task_info.p + 4 * ch.t0, out, plan->max_iter, plan->min_n, plan->max_n, part,
                           plan->flags, ev, plan->eig_order.p + ch.t0, ops);
        else
            launch_crossed(st, ni, nt, plan->insts2d.p + ch.i0, plan->tasks.p + ch.t0, plan->stacks.p,
                           plan->layouts.p, ws, plan->task_info.p + 4 * ch.t0, out, plan->max_iter, plan->min_n,
                           plan->max_n, part, plan->flags, ev, plan->eig_order.p + ch.t0, ops);
        hip_check(hipGetLastError(), "kernel launch");
    };
    for (size_t c = 0; c < nc; ++c) {
        if (nc > 1 && c >= 2) hip_check(hipStreamWaitEvent(front, plan->back_done[c - 2], 0), "hipStreamWaitEvent");
        launch(c, front, 1);
        if (nc > 1) {
            hip_check(hipEventRecord(plan->front_done[c], front), "hipEventRecord");
            hip_check(hipStreamWaitEvent(back, plan->front_done[c], 0), "hipStreamWaitEvent");
        }
        launch(c, back, 2);
        if (nc > 1) hip_check(hipEventRecord(plan->back_done[c], back), "hipEventRecord");
    }
    if (nc > 1) hip_check(hipStreamWaitEvent(front, plan->back_done[nc - 1], 0), "hipStreamWaitEvent");
}
}  // namespace

extern "C" {

const char* srcwa_last_error(void) { return g_error.c_str(); }
const char* srcwa_version(void) { return "srcwa-amd 0.1 (gfx950)"; }

int srcwa_context_create(int device, srcwa_context** out) {
    return guarded([&] {
        if (!out) invalid("srcwa_context_create: null output");
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device < 0 || device >= count) invalid("srcwa_context_create: no such device");
        hip_check(hipSetDevice(device), "hipSetDevice");
        // the no-allocation entry point's scratch array, located once per device (dense.hip)
        hip_check(static_cast<hipError_t>(dense_scratch_setup(device)), "hipGetSymbolAddress (scratch)");
        auto* c = new srcwa_context;
        c->device = device;
        const hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            hip_check(e, "hipStreamCreate");
        }
        *out = c;
    });
}

void srcwa_context_destroy(srcwa_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    for (hipStream_t st : {ctx->stream, ctx->back})
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    delete ctx;
}

int srcwa_context_sync(srcwa_context* ctx) {
    return guarded([&] {
        if (!ctx) invalid("null context");
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int srcwa_pool_create(int device, int64_t bytes, srcwa_pool** out) {
    return guarded([&] {
        if (!out || bytes <= 0) invalid("srcwa_pool_create: null output or non-positive size");
        hip_check(hipSetDevice(device), "hipSetDevice");
        auto pool = std::unique_ptr<srcwa_pool>(new srcwa_pool);
        pool->device = device;
        pool->bytes = srcwa_pool::round(static_cast<size_t>(bytes));
        hip_check(hipMalloc(reinterpret_cast<void**>(&pool->base), pool->bytes), "hipMalloc (pool)");
        pool->free_blocks[0] = pool->bytes;
        *out = pool.release();
    });
}

int srcwa_pool_destroy(srcwa_pool* pool) {
    return guarded([&] {
        if (!pool) return;
        if (pool->used != 0) invalid("srcwa_pool_destroy: plans still hold memory of this pool");
        hip_check(hipSetDevice(pool->device), "hipSetDevice");
        hip_check(hipFree(pool->base), "hipFree (pool)");
        delete pool;
    });
}

int srcwa_context_set_plan_flags(srcwa_context* ctx, int flags) {
    return guarded([&] {
        if (!ctx) invalid("srcwa_context_set_plan_flags: null context");
        if (flags & ~int(SRCWA_PLAN_DIAGNOSTICS | SRCWA_PLAN_MIXED | SRCWA_PLAN_POINT_OPS | SRCWA_PLAN_FIELDS))
            invalid("srcwa_context_set_plan_flags: unknown flag");
        if ((flags & SRCWA_PLAN_MIXED) && !SRCWA_MIXED_PRECISION) throw Failure{SRCWA_EUNSUPPORTED, kMixedOff};
        ctx->plan_flags = flags;
    });
}

int64_t srcwa_pool_free_bytes(srcwa_pool* pool) {
    if (!pool) return -1;
    std::lock_guard<std::mutex> g(pool->lock);
    return static_cast<int64_t>(pool->bytes - pool->used);
}

int srcwa_context_set_pool(srcwa_context* ctx, srcwa_pool* pool) {
    return guarded([&] {
        if (!ctx) invalid("srcwa_context_set_pool: null context");
        if (pool && pool->device != ctx->device) invalid("srcwa_context_set_pool: pool and context are on different devices");
        ctx->pool = pool;
    });
}

static int plan1d_create_impl(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                        int n_problems, const int* stack_index, const double* wavelength, const double* theta_deg,
                        const int* M, const int* pol, int64_t max_workspace_bytes, srcwa_plan** out, int64_t* need) {
    return guarded([&] {
        if (!ctx || !out) invalid("srcwa_plan1d_create: null context or output");
        if (n_stacks < 1 || n_problems < 1 || !stacks || !stack_offsets || !stack_index || !wavelength ||
            !theta_deg || !M || !pol)
            invalid("srcwa_plan1d_create: empty or null inputs");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        ParsedStacks ps = parse_all(n_stacks, stacks, stack_offsets);
        std::vector<std::vector<LayerInfo>>& infos = ps.infos;
        std::vector<double>& h_stacks = ps.words;

        auto plan = std::unique_ptr<srcwa_plan>(new srcwa_plan);
        plan->bind_pool(ctx->pool);
        plan->ctx = ctx;
        plan->kind = 1;
        plan->n_inst = n_problems;
        plan->flags = ctx->plan_flags;
        const bool diag = plan->flags & PLAN_DIAGNOSTICS;
        const bool fields = plan->flags & PLAN_FIELDS;
        const int64_t budget = max_workspace_bytes > 0 ? max_workspace_bytes : (int64_t(64) << 30);
        const int64_t budget_words = budget / int64_t(sizeof(cd));
        int64_t field_words = 0;

        std::vector<Inst1D> h_insts(n_problems);
        std::vector<LayerTask> h_tasks;
        std::vector<int64_t> h_layouts;
        int64_t orders = 0, chunk_ws = 0, max_ws = 0;
        PlanChunk cur{0, 0, 0, 0};
        OpSharing sharing;
        sharing.enabled = !(plan->flags & (PLAN_DIAGNOSTICS | PLAN_POINT_OPS));
        for (int i = 0; i < n_problems; ++i) {
            const int s = stack_index[i];
            if (s < 0 || s >= n_stacks) invalid("rcwa: stack index out of range");
            const double* st = stacks + stack_offsets[s];
            check_setup(st[0], wavelength[i], st[1], st[2], theta_deg[i], M[i], pol[i]);
            check_index(st[3], st[4]);
            const std::vector<LayerInfo>& li = infos[s];
            const int L = static_cast<int>(li.size());
            const int N = 2 * M[i] + 1;
            if (N > EIG_MAX_N_HOST) throw Failure{SRCWA_EUNSUPPORTED, "rcwa: truncation order above this build's limit"};
            std::vector<bool> general(L + 2, false);
            for (int j = 0; j < L; ++j) general[j + 1] = li[j].grating;
            const PlanarCost cost = planar_cost(N, general, pol[i]);
            plan->flops[0] += cost.flops_eig;
            plan->flops[1] += cost.flops_other;
            // layout (offsets relative to the instance; rebased onto the chunk below)
            int64_t w = 0;
            std::vector<int64_t> lay(layout_size(L), -1);
            lay[LAYOUT_KX] = w;
            w += align16(N);
            for (int r = 0; r < L + 2; ++r) {
                lay[layout_q(r)] = w;
                w += align16(N);
                lay[layout_p(r)] = w;
                w += align16(N);
            }
            // with diagnostics each grating layer keeps its matrix and a solve scratch (2 N^2), and the
            // recursion's solves copies of A, B and the residual (3 N^2)
            const int64_t NN1 = int64_t(N) * N;
            for (int j = 0; j < L; ++j)
                if (li[j].grating) {
                    lay[layout_layer(L, j)] = w;
                    w += align16(slot_words(N) + (diag ? 2 * NN1 : 0) + fft_scratch_words(li[j], M[i]));
                }
            const int64_t ws_lo = w;
            w += cost.bump_words + (diag ? 3 * NN1 + 64 : 0) + ((plan->flags & PLAN_MIXED) ? mixed_solve_words(N, N) : 0);
            // near fields: the reflection below every layer is kept, and one more vector
            if (fields) w += int64_t(L) * align16(NN1) + align16(N) + 64;
            // admission into the current chunk by workspace footprint
            if (i > cur.i0 && chunk_ws + w > budget_words) {
                cur.i1 = i;
                cur.t1 = static_cast<int>(h_tasks.size());
                plan->chunks.push_back(cur);
                max_ws = std::max(max_ws, chunk_ws);
                cur = PlanChunk{i, i, cur.t1, cur.t1};
                chunk_ws = 0;
            }
            const int first_task = static_cast<int>(h_tasks.size());
            for (int j = 0; j < L; ++j)
                if (li[j].grating) {
                    h_tasks.push_back(LayerTask{i - cur.i0, j});
                    // TE: [[eps]]; TM: also [[1/eps]] and both LU factorizations (rcwa1d.hpp:432-468)
                    std::vector<uint64_t> key;
                    for (double v : {st[0], double(M[i]), double(pol[i])}) OpSharing::push(key, v);
                    for (int64_t k = 0; k < li[j].rec_words; ++k) OpSharing::push(key, st[li[j].rec + k]);
                    sharing.add(key, i, j, planar_ops_words(N, pol[i]) + fft_scratch_words(li[j], M[i]));
                }
            const int64_t layout_off = static_cast<int64_t>(h_layouts.size());
            for (int64_t v : lay) h_layouts.push_back(v >= 0 ? v + chunk_ws : -1);
            Inst1D& in = h_insts[i];
            in.M = M[i];
            in.N = N;
            in.flags = plan->flags;
            plan->max_n = std::max(plan->max_n, N);
            plan->min_n = std::min(plan->min_n, N);
            in.pol = pol[i];
            in.L = L;
            in.first_task = first_task - cur.t0;
            in.n_tasks = static_cast<int>(h_tasks.size()) - first_task;
            in.wavelength = wavelength[i];
            in.theta_deg = theta_deg[i];
            in.phi_deg = 0.0;
            in.stack_off = stack_offsets[s];
            in.layout_off = layout_off;
            in.out_off = orders;
            in.ws_lo = chunk_ws + ws_lo;
            in.ws_hi = chunk_ws + w;
            in.field_off = field_words;
            if (fields) {
                field_words += 2 * int64_t(L) * N;
                plan->h_layers.push_back(li);
            }
            chunk_ws += w;
            orders += N;
        }
        cur.i1 = n_problems;
        cur.t1 = static_cast<int>(h_tasks.size());
        plan->chunks.push_back(cur);
        max_ws = std::max(max_ws, chunk_ws);
        // the samplers read the modes where the solve left them, so a field plan runs as one chunk
        if (fields && plan->chunks.size() > 1)
            invalid("rcwa: a plan with near fields must fit one workspace chunk (raise max_workspace_bytes)");
        plan->n_tasks = static_cast<int>(h_tasks.size());
        plan->total_orders = orders;
        std::vector<LayerOps> h_ops;
        const int64_t op_words = sharing.assign(h_tasks, h_ops);

        if (need) {
            const auto b = [](size_t x) { return static_cast<int64_t>(srcwa_pool::round(x)); };
            *need = b(h_insts.size() * sizeof(Inst1D)) + b(h_tasks.size() * sizeof(LayerTask)) +
                    b(h_stacks.size() * sizeof(double)) + b(h_layouts.size() * sizeof(int64_t)) +
                    b((plan->chunks.size() > 1 ? 2 : 1) * max_ws * sizeof(cd)) + b(std::max(1, plan->n_tasks) * sizeof(int)) +
                    2 * b(2 * orders * sizeof(double)) + 2 * b(orders * sizeof(double)) + 2 * b(n_problems * sizeof(int)) +
                    diag_bytes(n_problems, plan->n_tasks) + ops_bytes(h_ops, op_words) +
                    (fields ? b(std::max<int64_t>(1, field_words) * sizeof(cd)) + b(n_problems * sizeof(int)) : 0);
            return;
        }
        hipStream_t st = ctx->stream;
        plan->insts.upload(h_insts, st);
        if (fields) {
            plan->fields.alloc(static_cast<size_t>(std::max<int64_t>(1, field_words)));
            plan->field_ok.alloc(n_problems);
            plan->h_insts = h_insts;
            plan->h_stack_words = h_stacks;
        }
        plan->tasks.upload(h_tasks, st);
        plan->upload_ops(h_ops, op_words, st);
        plan->stacks.upload(h_stacks, st);
        plan->layouts.upload(h_layouts, st);
        plan->region_words = max_ws;
        plan->ws.alloc(static_cast<size_t>(plan->chunks.size() > 1 ? 2 * max_ws : max_ws));
        plan->task_info.alloc(std::max(1, plan->n_tasks));
        plan->R.alloc(2 * orders);
        plan->T.alloc(2 * orders);
        plan->DEr.alloc(orders);
        plan->DEt.alloc(orders);
        plan->status.alloc(n_problems);
        plan->eig_info.alloc(n_problems);
        plan->alloc_diag();
        plan->make_events();
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        *out = plan.release();
    });
}

int srcwa_plan1d_create(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                        int n_problems, const int* stack_index, const double* wavelength, const double* theta_deg,
                        const int* M, const int* pol, int64_t max_workspace_bytes, srcwa_plan** out) {
    return plan1d_create_impl(ctx, n_stacks, stacks, stack_offsets, n_problems, stack_index, wavelength, theta_deg,
           M, pol, max_workspace_bytes, out, nullptr);
}

int srcwa_plan1d_bytes(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                        int n_problems, const int* stack_index, const double* wavelength, const double* theta_deg,
                        const int* M, const int* pol, int64_t max_workspace_bytes, int64_t* bytes) {
    if (!bytes) {
        g_error = "null output";
        return SRCWA_EINVAL;
    }
    srcwa_plan* unused = nullptr;
    return plan1d_create_impl(ctx, n_stacks, stacks, stack_offsets, n_problems, stack_index, wavelength, theta_deg,
           M, pol, max_workspace_bytes, &unused, bytes);
}

static int plan_conical_create_impl(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                              int n_problems, const int* stack_index, const double* wavelength,
                              const double* theta_deg, const double* phi_deg, const int* M,
                              int64_t max_workspace_bytes, srcwa_plan** out, int64_t* need) {
    return guarded([&] {
        if (!ctx || !out) invalid("srcwa_plan_conical_create: null context or output");
        if (n_stacks < 1 || n_problems < 1 || !stacks || !stack_offsets || !stack_index || !wavelength ||
            !theta_deg || !phi_deg || !M)
            invalid("srcwa_plan_conical_create: empty or null inputs");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        ParsedStacks ps = parse_all(n_stacks, stacks, stack_offsets);
        auto plan = std::unique_ptr<srcwa_plan>(new srcwa_plan);
        plan->bind_pool(ctx->pool);
        plan->ctx = ctx;
        plan->kind = 2;
        plan->n_inst = n_problems;
        plan->flags = ctx->plan_flags;
        const bool diag = plan->flags & PLAN_DIAGNOSTICS;
        const bool fields = plan->flags & PLAN_FIELDS;
        int64_t field_words = 0;
        const int64_t budget = max_workspace_bytes > 0 ? max_workspace_bytes : (int64_t(64) << 30);
        const int64_t budget_words = budget / int64_t(sizeof(cd));
        std::vector<Inst1D> h_insts(n_problems);
        std::vector<LayerTask> h_tasks;
        std::vector<int64_t> h_layouts;
        int64_t orders = 0, chunk_ws = 0, max_ws = 0;
        PlanChunk cur{0, 0, 0, 0};
        OpSharing sharing;
        // a field plan keeps every layer's [[eps]]^-1 in its own slots, where the sampler reads it
        sharing.enabled = !(plan->flags & (PLAN_DIAGNOSTICS | PLAN_POINT_OPS | PLAN_FIELDS));
        for (int i = 0; i < n_problems; ++i) {
            const int s = stack_index[i];
            if (s < 0 || s >= n_stacks) invalid("rcwa: stack index out of range");
            const double* st = stacks + stack_offsets[s];
            // rcwa1d_conical.hpp:686-693
            check_setup(st[0], wavelength[i], st[1], st[2], theta_deg[i], M[i], 0);
            if (!std::isfinite(phi_deg[i])) invalid("rcwa: azimuth angle must be finite");
            check_index(st[3], st[4]);
            if (st[4] != 0.0) invalid("rcwa: conical incidence needs a real index in the exit medium");
            const std::vector<LayerInfo>& all = ps.infos[s];
            std::vector<int> eff;                           // layers with d > 0
            for (int j = 0; j < static_cast<int>(all.size()); ++j)
                if (all[j].d > 0.0) eff.push_back(j);
            const int L = static_cast<int>(eff.size());
            const int N = 2 * M[i] + 1;
            if (N > EIG_MAX_N_HOST) throw Failure{SRCWA_EUNSUPPORTED, "rcwa: truncation order above this build's limit"};
            std::vector<bool> general(L + 2, false);
            for (int j = 0; j < L; ++j) general[j + 1] = all[eff[j]].grating && !all[eff[j]].uniform_values;
            PlanarCost cost = planar_cost(2 * N, general, 0, true);
            // degenerate layers borrow four (2N)^2 scratch matrices each from this region
            int n_general = 0;
            for (int j = 1; j <= L; ++j) n_general += general[j];
            cost.bump_words = std::max<int64_t>(cost.bump_words, 16 * int64_t(N) * N * n_general + 64);
            plan->flops[0] += cost.flops_eig;
            plan->flops[1] += cost.flops_other;
            int64_t w = 0;
            std::vector<int64_t> lay(conical_layout_size(L), -1);
            lay[0] = w;
            w += align16(6 * N + 1);
            for (int r = 0; r < L + 2; ++r) {
                lay[conical_q(r)] = w;
                w += align16(2 * N);
                const bool analytic = r == 0 || r == L + 1 || !general[r];
                if (analytic) {
                    lay[conical_w(r)] = w;
                    w += align16(2 * N);
                    lay[conical_v(r)] = w;
                    w += align16(2 * N);
                }
            }
            // with diagnostics: copies of A and B C plus two solve scratch matrices per layer (4 N^2), and
            // the recursion's copies of A, B and the residual at order 2N (3 (2N)^2)
            const int64_t NN1 = int64_t(N) * N;
            for (int j = 0; j < L; ++j)
                if (general[j + 1]) {
                    lay[conical_slot(L, j)] = w;
                    w += align16(conical_slot_words(N) + (diag ? 4 * NN1 : 0) + fft_scratch_words(all[eff[j]], M[i]) +
                                 (fields ? 4 * NN1 : 0));      // field plans keep a Schur generator here
                }
            const int64_t ws_lo = w;
            w += cost.bump_words + (diag ? 12 * NN1 + 64 : 0) +
                 ((plan->flags & PLAN_MIXED) ? mixed_solve_words(2 * N, 2 * N) : 0);
            // near fields: the reflection below every layer (order 2N) is kept, and one more 2N x 2 block
            if (fields) w += int64_t(L) * align16(4 * NN1) + align16(4 * int64_t(N)) + 64;
            if (i > cur.i0 && chunk_ws + w > budget_words) {
                cur.i1 = i;
                cur.t1 = static_cast<int>(h_tasks.size());
                plan->chunks.push_back(cur);
                max_ws = std::max(max_ws, chunk_ws);
                cur = PlanChunk{i, i, cur.t1, cur.t1};
                chunk_ws = 0;
            }
            const int first_task = static_cast<int>(h_tasks.size());
            for (int j = 0; j < L; ++j)
                if (general[j + 1]) {
                    h_tasks.push_back(LayerTask{i - cur.i0, j});
                    // [[eps]], [[eps]]^-1 and [[1/eps]]^-1 (rcwa1d_conical.hpp:302-336)
                    const LayerInfo& lj = all[eff[j]];
                    std::vector<uint64_t> key;
                    for (double v : {st[0], double(M[i])}) OpSharing::push(key, v);
                    for (int64_t k = 0; k < lj.rec_words; ++k) OpSharing::push(key, st[lj.rec + k]);
                    sharing.add(key, i, eff[j], conical_ops_words(N) + fft_scratch_words(lj, M[i]));
                }
            const int64_t layout_off = static_cast<int64_t>(h_layouts.size());
            for (int r = 0; r < conical_layout_size(L); ++r) {
                int64_t v = lay[r];
                if (r >= conical_orig(L, 0) && L > 0) v = eff[r - conical_orig(L, 0)];
                else if (v >= 0) v += chunk_ws;
                h_layouts.push_back(v);
            }
            Inst1D& in = h_insts[i];
            in.M = M[i];
            in.N = N;
            plan->max_n = std::max(plan->max_n, N);
            plan->min_n = std::min(plan->min_n, N);
            in.pol = 2;
            in.flags = plan->flags;
            in.L = L;
            in.first_task = first_task - cur.t0;
            in.n_tasks = static_cast<int>(h_tasks.size()) - first_task;
            in.wavelength = wavelength[i];
            in.theta_deg = theta_deg[i];
            in.phi_deg = phi_deg[i];
            in.stack_off = stack_offsets[s];
            in.layout_off = layout_off;
            in.out_off = orders;
            in.ws_lo = chunk_ws + ws_lo;
            in.ws_hi = chunk_ws + w;
            in.field_off = field_words;
            if (fields) {
                // per layer a+ and a- (2N x 2 each), then the outer responses up and down (2N x 2 each), then one
                // Schur-basis mark per layer
                field_words += 8 * int64_t(N) * (L + 1) + L;
                plan->h_layers.push_back(all);
            }
            chunk_ws += w;
            orders += N;
        }
        cur.i1 = n_problems;
        cur.t1 = static_cast<int>(h_tasks.size());
        plan->chunks.push_back(cur);
        max_ws = std::max(max_ws, chunk_ws);
        if (fields && plan->chunks.size() > 1)
            invalid("rcwa: a plan with near fields must fit one workspace chunk (raise max_workspace_bytes)");
        plan->n_tasks = static_cast<int>(h_tasks.size());
        plan->total_orders = orders;
        std::vector<LayerOps> h_ops;
        const int64_t op_words = sharing.assign(h_tasks, h_ops);
        if (need) {
            const auto b = [](size_t x) { return static_cast<int64_t>(srcwa_pool::round(x)); };
            *need = b(h_insts.size() * sizeof(Inst1D)) + b(h_tasks.size() * sizeof(LayerTask)) +
                    b(std::max<size_t>(1, h_tasks.size()) * sizeof(int)) + b(ps.words.size() * sizeof(double)) + b(h_layouts.size() * sizeof(int64_t)) +
                    b((plan->chunks.size() > 1 ? 2 : 1) * max_ws * sizeof(cd)) + b(4 * std::max(1, plan->n_tasks) * sizeof(int)) +
                    2 * b(8 * orders * sizeof(double)) + 2 * b(4 * orders * sizeof(double)) +
                    b(GEO_WORDS * orders * sizeof(double)) + 2 * b(n_problems * sizeof(int)) +
                    diag_bytes(n_problems, plan->n_tasks) + ops_bytes(h_ops, op_words) +
                    (fields ? b(std::max<int64_t>(1, field_words) * sizeof(cd)) + b(n_problems * sizeof(int)) : 0);
            return;
        }
        hipStream_t st = ctx->stream;
        plan->insts.upload(h_insts, st);
        if (fields) {
            plan->fields.alloc(static_cast<size_t>(std::max<int64_t>(1, field_words)));
            plan->field_ok.alloc(n_problems);
            plan->h_insts = h_insts;
            plan->h_stack_words = ps.words;
        }
        plan->tasks.upload(h_tasks, st);
        plan->upload_ops(h_ops, op_words, st);
        plan->eig_order.upload(longest_first(plan->chunks, h_tasks, h_insts), st);
        plan->stacks.upload(ps.words, st);
        plan->layouts.upload(h_layouts, st);
        plan->region_words = max_ws;
        plan->ws.alloc(static_cast<size_t>(plan->chunks.size() > 1 ? 2 * max_ws : max_ws));
        plan->task_info.alloc(4 * std::max(1, plan->n_tasks));
        plan->R.alloc(8 * orders);
        plan->T.alloc(8 * orders);
        plan->DEr.alloc(4 * orders);
        plan->geometry.alloc(GEO_WORDS * orders);
        plan->DEt.alloc(4 * orders);
        plan->status.alloc(n_problems);
        plan->eig_info.alloc(n_problems);
        plan->alloc_diag();
        plan->make_events();
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        *out = plan.release();
    });
}

int srcwa_plan_conical_create(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                              int n_problems, const int* stack_index, const double* wavelength,
                              const double* theta_deg, const double* phi_deg, const int* M,
                              int64_t max_workspace_bytes, srcwa_plan** out) {
    return plan_conical_create_impl(ctx, n_stacks, stacks, stack_offsets, n_problems, stack_index, wavelength,
           theta_deg, phi_deg, M, max_workspace_bytes, out, nullptr);
}

int srcwa_plan_conical_bytes(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                              int n_problems, const int* stack_index, const double* wavelength,
                              const double* theta_deg, const double* phi_deg, const int* M,
                              int64_t max_workspace_bytes, int64_t* bytes) {
    if (!bytes) {
        g_error = "null output";
        return SRCWA_EINVAL;
    }
    srcwa_plan* unused = nullptr;
    return plan_conical_create_impl(ctx, n_stacks, stacks, stack_offsets, n_problems, stack_index, wavelength,
           theta_deg, phi_deg, M, max_workspace_bytes, &unused, bytes);
}

int srcwa_plan_execute(srcwa_plan* plan) {
    return guarded([&] {
        if (!plan) invalid("null plan");
        hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
        hipStream_t front = plan->ctx->stream;
        if (plan->graph_exec) {
            hip_check(hipEventRecord(plan->ev.front(), front), "hipEventRecord");
            hip_check(hipGraphLaunch(plan->graph_exec, front), "hipGraphLaunch");
        } else {
            enqueue(plan, true);
        }
        hip_check(hipEventRecord(plan->ev.back(), front), "hipEventRecord");
        plan->timed = true;
        plan->executed = true;
    });
}

int srcwa_plan_set_graph(srcwa_plan* plan, int enable) {
    return guarded([&] {
        if (!plan) invalid("null plan");
        hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
        if (plan->graph_exec) {
            hip_check(hipGraphExecDestroy(plan->graph_exec), "hipGraphExecDestroy");
            plan->graph_exec = nullptr;
        }
        if (plan->graph) {
            hip_check(hipGraphDestroy(plan->graph), "hipGraphDestroy");
            plan->graph = nullptr;
        }
        if (!enable) return;
        if (plan->chunks.size() > 1)
            throw Failure{SRCWA_EUNSUPPORTED, "srcwa_plan_set_graph: graph replay needs a plan of one chunk"};
        hipStream_t st = plan->ctx->stream;
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        try {
            enqueue(plan, false);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(st, &plan->graph), "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&plan->graph_exec, plan->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    });
}

int srcwa_plan_set_max_iter(srcwa_plan* plan, int max_iter) {
    if (!plan) {
        g_error = "null plan";
        return SRCWA_EINVAL;
    }
    if (max_iter < 1 || max_iter > 100000) {
        g_error = "srcwa_plan_set_max_iter: max_iter must lie in 1..100000";
        return SRCWA_EINVAL;
    }
    plan->max_iter = max_iter;
    // a captured graph holds the old limit in its kernel arguments: capture again
    return plan->graph_exec ? srcwa_plan_set_graph(plan, 1) : SRCWA_OK;
}

int srcwa_plan_sync(srcwa_plan* plan) {
    return guarded([&] {
        if (!plan) invalid("null plan");
        hip_check(hipStreamSynchronize(plan->ctx->stream), "hipStreamSynchronize");
    });
}

int srcwa_plan_poll(srcwa_plan* plan) {
    if (!plan) {
        g_error = "null plan";
        return SRCWA_EINVAL;
    }
    if (plan->ev.empty()) return 1;
    const hipError_t e = hipEventQuery(plan->ev.back());
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    g_error = std::string("hipEventQuery: ") + hipGetErrorString(e);
    return SRCWA_EDEVICE;
}

int srcwa_plan_results(srcwa_plan* plan, double* R, double* T, double* DEr, double* DEt, int* status, int* eig_info) {
    return guarded([&] {
        if (!plan) invalid("null plan");
        hipStream_t s = plan->ctx->stream;
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        const int64_t o = plan->total_orders * (plan->kind == 2 ? 4 : 1);
        if (plan->kind == 3) {
            // crossed gratings: R = [Rs, Rp, Ex_r, Ey_r], T = [Ts, Tp, Ex_t, Ey_t] per instance (4N complex)
            if (R) copy_out(R, plan->R.p, 8 * o * sizeof(double), hipMemcpyDeviceToHost, "D2H", plan);
            if (T) copy_out(T, plan->T.p, 8 * o * sizeof(double), hipMemcpyDeviceToHost, "D2H", plan);
            R = T = nullptr;
        }
        if (R) copy_out(R, plan->R.p, 2 * o * sizeof(double), hipMemcpyDeviceToHost, "D2H", plan);
        if (T) copy_out(T, plan->T.p, 2 * o * sizeof(double), hipMemcpyDeviceToHost, "D2H", plan);
        if (DEr) copy_out(DEr, plan->DEr.p, o * sizeof(double), hipMemcpyDeviceToHost, "D2H", plan);
        if (DEt) copy_out(DEt, plan->DEt.p, o * sizeof(double), hipMemcpyDeviceToHost, "D2H", plan);
        if (status) copy_out(status, plan->status.p, plan->n_inst * sizeof(int), hipMemcpyDeviceToHost, "D2H", plan);
        if (eig_info)
            copy_out(eig_info, plan->eig_info.p, plan->n_inst * sizeof(int), hipMemcpyDeviceToHost, "D2H", plan);
    });
}

// ---- near fields of planar and conical plans -----------------------------------------------------
namespace {
// The capture of one problem is usable: the plan keeps fields, has run, the solve succeeded and every
// captured mode and amplitude is finite (rcwa1d.hpp:1057-1066, rcwa1d_conical.hpp:957-970).
bool field_ready(srcwa_plan* plan, int problem) {
    int st = 1, ok = 0;
    copy_out(&st, plan->status.p + problem, sizeof(int), hipMemcpyDeviceToHost, "D2H", plan);
    copy_out(&ok, plan->field_ok.p + problem, sizeof(int), hipMemcpyDeviceToHost, "D2H", plan);
    return st == ST_OK && ok == 1;
}

// kind: 1 planar, 2 conical, 0 either
void field_plan_checks(srcwa_plan* plan, int problem, const char* what, int kind) {
    if (!plan) invalid(std::string(what) + ": null plan");
    if (plan->kind > 2 || (kind && plan->kind != kind) || !(plan->flags & PLAN_FIELDS))
        invalid(std::string(what) + ": the plan keeps no near fields of this kind (create it with SRCWA_PLAN_FIELDS)");
    if (problem < 0 || problem >= plan->n_inst) invalid(std::string(what) + ": problem index out of range");
    if (!plan->executed) invalid(std::string(what) + ": the plan has not been executed");
    hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
    for (hipStream_t st : {plan->ctx->stream, plan->ctx->back})
        if (st) hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

// normalized_jones (rcwa_fields.hpp:218-230): scaled by its largest component, then to unit norm
void normalize_jones(const double* jones, cd& j0, cd& j1) {
    if (!jones) invalid("rcwa: Jones input must be finite and nonzero");
    std::complex<double> v[2] = {{jones[0], jones[1]}, {jones[2], jones[3]}};
    double scale = 0.0;
    for (const auto& c : v) {
        if (!std::isfinite(c.real()) || !std::isfinite(c.imag())) invalid("rcwa: Jones input must be finite and nonzero");
        scale = std::max(scale, std::max(std::abs(c.real()), std::abs(c.imag())));
    }
    if (scale == 0.0) invalid("rcwa: Jones input must be finite and nonzero");
    for (auto& c : v) c /= scale;
    const double norm = std::sqrt(std::norm(v[0]) + std::norm(v[1]));
    for (auto& c : v) c /= norm;
    j0 = cd{v[0].real(), v[0].imag()};
    j1 = cd{v[1].real(), v[1].imag()};
}

// Harmonic fields of problem `problem` at nz depths into the device buffer H (6 x N x nz complex), with the
// reference's depth groups (field_depth_groups, rcwa_fields.hpp:253-279): z < 0 the incidence side, else the
// exit side unless some layer with modes has its bottom at or below z + 1e-12 (the first such layer; an
// interface belongs to the layer above).  Planar plans keep every layer; conical plans only the layers of
// positive thickness (copy_field_geometry with skip_zero, rcwa_fields.hpp:403-434).
void sample_fields(srcwa_plan* plan, int problem, int nz, const double* zs, DevBuf<cd>& H, cd j0, cd j1) {
    const Inst1D& in = plan->h_insts[problem];
    const std::vector<LayerInfo>& li = plan->h_layers[problem];
    const bool conical = plan->kind == 2;
    const int N = in.N, all = static_cast<int>(li.size());
    for (int iz = 0; iz < nz; ++iz)
        if (!std::isfinite(zs[iz])) invalid("rcwa: field depths must be finite");
    // the layers with modes: planar all, conical those with d > 0 (region index = position in this list)
    std::vector<int> idx;
    std::vector<double> z_top, z_bot;
    double z = 0.0;
    bool finite_layer = false;
    for (int j = 0; j < all; ++j) {
        const double top = z;
        z += li[j].d;
        if (conical && !(li[j].d > 0.0)) continue;
        idx.push_back(j);
        z_top.push_back(top);
        z_bot.push_back(z);
        finite_layer = finite_layer || li[j].d > 0.0;
    }
    const int L = static_cast<int>(idx.size());
    const double z_bottom = z;
    std::vector<std::vector<int>> members(L + 2);   // 0 incidence, 1 exit, 2 + j layer j
    for (int iz = 0; iz < nz; ++iz) {
        const double zi = zs[iz];
        int grp = zi < 0.0 ? 0 : 1;
        if (zi >= 0.0 && zi <= z_bottom && finite_layer)
            for (int j = 0; j < L; ++j)
                if (zi <= z_bot[j] + 1e-12) {
                    grp = 2 + j;
                    break;
                }
        members[grp].push_back(iz);
    }
    std::vector<FieldGroup1D> groups;
    std::vector<int> cols;
    int64_t scr = 0;
    const int64_t NN = int64_t(N) * N;
    for (int gi = 0; gi < L + 2; ++gi) {
        if (members[gi].empty()) continue;
        FieldGroup1D g{};
        g.region = gi == 0 ? -1 : gi == 1 ? -2 : gi - 2;
        g.first = static_cast<int>(cols.size());
        g.count = static_cast<int>(members[gi].size());
        g.tm_grating = !conical && in.pol == 1 && g.region >= 0 && li[idx[g.region]].grating;
        g.z_top = g.region >= 0 ? z_top[g.region] : 0.0;
        g.d = g.region >= 0 ? li[idx[g.region]].d : 0.0;
        g.z_bottom = z_bottom;
        g.scratch = scr;
        const int64_t nc = int64_t(N) * g.count;
        if (conical) {
            scr += 4 * align16(2 * nc) + 2 * align16(nc) + 16;
            // a layer group may sample a Schur-basis layer: its anchors and four (2N)^2 exponential buffers
            if (g.region >= 0) scr += 4 * align16(4 * NN) + align16(6 * int64_t(N)) + 64;
        } else {
            scr += 5 * align16(nc);
            if (g.tm_grating)
                scr += align16(NN) + align16(N) + align16(4 * in.M + 1) + fft_scratch_words(li[idx[g.region]], in.M) + 16;
        }
        cols.insert(cols.end(), members[gi].begin(), members[gi].end());
        groups.push_back(g);
    }
    hipStream_t st = plan->ctx->stream;
    H.alloc(static_cast<size_t>(6 * int64_t(N) * nz));
    DevBuf<FieldGroup1D> d_groups;
    DevBuf<int> d_cols;
    DevBuf<double> d_zs;
    DevBuf<cd> d_scr;
    d_groups.pool = d_cols.pool = plan->ctx->pool;      // a pooled context takes these from its pool too
    d_zs.pool = plan->ctx->pool;
    d_scr.pool = plan->ctx->pool;
    const std::vector<double> h_zs(zs, zs + nz);     // lives until the launch below has been synchronized
    d_groups.upload(groups, st);
    d_cols.upload(cols, st);
    d_zs.upload(h_zs, st);
    d_scr.alloc(static_cast<size_t>(std::max<int64_t>(16, scr)));
    if (conical) {
        launch_fields_conical(st, static_cast<int>(groups.size()), plan->insts.p + problem, plan->stacks.p,
                              plan->layouts.p, plan->ws.p, plan->fields.p, d_groups.p, d_cols.p, d_zs.p, nz, j0, j1,
                              H.p, d_scr.p);
    } else {
        const int64_t oo = std::accumulate(plan->h_insts.begin(), plan->h_insts.begin() + problem, int64_t(0),
                                           [](int64_t a, const Inst1D& x) { return a + x.N; });
        launch_fields1d(st, static_cast<int>(groups.size()), plan->insts.p + problem, plan->stacks.p, plan->layouts.p,
                        plan->ws.p, plan->fields.p, reinterpret_cast<const cd*>(plan->R.p) + oo,
                        reinterpret_cast<const cd*>(plan->T.p) + oo, d_groups.p, d_cols.p, d_zs.p, nz, H.p, d_scr.p);
    }
    hip_check(hipGetLastError(), "kernel launch");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

bool all_finite(const std::vector<double>& v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

const char* unavailable_msg(const srcwa_plan* plan) {
    return plan->kind == 2 ? "rcwa: conical field context is unavailable" : "rcwa: planar field context is unavailable";
}
const char* nonfinite_msg(const srcwa_plan* plan) {
    return plan->kind == 2 ? "rcwa: conical field propagation produced nonfinite values"
                           : "rcwa: planar field propagation produced nonfinite values";
}

// field_at of either kind; jones: conical only
void field_at_impl(srcwa_plan* plan, int problem, int nz, const double* zs, const double* jones, double* out,
                   const char* what, int kind) {
    field_plan_checks(plan, problem, what, kind);
    if (!field_ready(plan, problem)) throw Failure{SRCWA_EUNAVAILABLE, unavailable_msg(plan)};
    cd j0{1.0, 0.0}, j1{0.0, 0.0};
    if (kind == 2) normalize_jones(jones, j0, j1);
    if (nz < 0 || (nz > 0 && (!zs || !out))) invalid(std::string(what) + ": bad depth arguments");
    if (nz == 0) return;
    DevBuf<cd> H;
    H.pool = plan->ctx->pool;
    sample_fields(plan, problem, nz, zs, H, j0, j1);
    const size_t words = 12 * size_t(plan->h_insts[problem].N) * nz;
    std::vector<double> h(words);
    copy_out(h.data(), H.p, words * sizeof(double), hipMemcpyDeviceToHost, "D2H", plan);
    if (!all_finite(h)) throw Failure{SRCWA_ENUMERICAL, nonfinite_msg(plan)};
    std::memcpy(out, h.data(), words * sizeof(double));
}

void field_map_impl(srcwa_plan* plan, int problem, int nz, const double* zs, int nx, double y, const double* jones,
                    double* out, const char* what, int kind) {
    if (nx <= 0) invalid("rcwa: field nx must be positive");
    if (!std::isfinite(y)) invalid("rcwa: field y must be finite");
    field_plan_checks(plan, problem, what, kind);
    if (!field_ready(plan, problem)) throw Failure{SRCWA_EUNAVAILABLE, unavailable_msg(plan)};
    cd j0{1.0, 0.0}, j1{0.0, 0.0};
    if (kind == 2) normalize_jones(jones, j0, j1);
    if (nz < 0 || (nz > 0 && (!zs || !out))) invalid(std::string(what) + ": bad depth arguments");
    if (nz == 0) return;
    DevBuf<cd> H;
    H.pool = plan->ctx->pool;
    sample_fields(plan, problem, nz, zs, H, j0, j1);
    const Inst1D& in = plan->h_insts[problem];
    const int N = in.N;
    std::vector<double> h(12 * size_t(N) * nz);
    copy_out(h.data(), H.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, "D2H", plan);
    if (!all_finite(h)) throw Failure{SRCWA_ENUMERICAL, nonfinite_msg(plan)};
    // the transverse phase ky y of conical solves (the field metadata's ky = k0 nI sin(theta) sin(phi))
    double kyy = 0.0;
    if (kind == 2) {
        const double* st = plan->h_stack_words.data() + in.stack_off;
        const double k0 = 2.0 * kPi / in.wavelength;
        const double ky = st[1] * std::sin(in.theta_deg * kPi / 180.0) * std::sin(in.phi_deg * kPi / 180.0);
        kyy = (k0 * ky) * y;
    }
    DevBuf<cd> G;
    G.pool = plan->ctx->pool;
    G.alloc(static_cast<size_t>(6 * int64_t(nx) * nz));
    launch_fieldmap1d(plan->ctx->stream, plan->insts.p + problem, plan->stacks.p, plan->layouts.p, plan->ws.p, H.p,
                      nz, nx, kyy, G.p);
    hip_check(hipGetLastError(), "kernel launch");
    std::vector<double> g(12 * size_t(nx) * nz);
    copy_out(g.data(), G.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, "D2H", plan);
    if (!all_finite(g)) throw Failure{SRCWA_ENUMERICAL, "rcwa: field reconstruction produced nonfinite values"};
    std::memcpy(out, g.data(), g.size() * sizeof(double));
}
}  // namespace

int srcwa_plan1d_field_available(srcwa_plan* plan, int problem, int* available) {
    return guarded([&] {
        field_plan_checks(plan, problem, "srcwa_plan1d_field_available", 0);
        if (!available) invalid("srcwa_plan1d_field_available: null output");
        *available = field_ready(plan, problem) ? 1 : 0;
    });
}

int srcwa_plan1d_field_at(srcwa_plan* plan, int problem, int nz, const double* zs, double* out) {
    return guarded([&] { field_at_impl(plan, problem, nz, zs, nullptr, out, "srcwa_plan1d_field_at", 1); });
}

int srcwa_plan1d_field_map(srcwa_plan* plan, int problem, int nz, const double* zs, int nx, double y, double* out) {
    return guarded([&] { field_map_impl(plan, problem, nz, zs, nx, y, nullptr, out, "srcwa_plan1d_field_map", 1); });
}

int srcwa_plan1d_field_layers(srcwa_plan* plan, int problem, int* zero_modes) {
    return guarded([&] {
        field_plan_checks(plan, problem, "srcwa_plan1d_field_layers", 1);
        if (!zero_modes) invalid("srcwa_plan1d_field_layers: null output");
        if (!field_ready(plan, problem)) throw Failure{SRCWA_EUNAVAILABLE, unavailable_msg(plan)};
        const Inst1D& in = plan->h_insts[problem];
        const int L = in.L, N = in.N;
        std::vector<int64_t> lay(static_cast<size_t>(layout_size(L)));
        copy_out(lay.data(), plan->layouts.p + in.layout_off, lay.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                 "D2H", plan);
        std::vector<cd> q(static_cast<size_t>(N));
        for (int j = 0; j < L; ++j) {
            copy_out(q.data(), plan->ws.p + lay[layout_q(j + 1)], N * sizeof(cd), hipMemcpyDeviceToHost, "D2H", plan);
            int zeros = 0;
            for (const cd& v : q) zeros += v.x == 0.0 && v.y == 0.0;
            zero_modes[j] = zeros;
        }
    });
}

int srcwa_plan_conical_field_layers(srcwa_plan* plan, int problem, int* schur) {
    return guarded([&] {
        field_plan_checks(plan, problem, "srcwa_plan_conical_field_layers", 2);
        if (!schur) invalid("srcwa_plan_conical_field_layers: null output");
        if (!field_ready(plan, problem)) throw Failure{SRCWA_EUNAVAILABLE, unavailable_msg(plan)};
        const Inst1D& in = plan->h_insts[problem];
        const int L = in.L, n = 2 * in.N;
        std::vector<cd> marks(static_cast<size_t>(std::max(1, L)));
        if (L > 0)
            copy_out(marks.data(), plan->fields.p + in.field_off + 4 * int64_t(L + 1) * n, L * sizeof(cd),
                     hipMemcpyDeviceToHost, "D2H", plan);
        for (int j = 0; j < L; ++j) schur[j] = marks[j].x != 0.0 ? 1 : 0;
    });
}

int srcwa_plan_conical_field_at(srcwa_plan* plan, int problem, int nz, const double* zs, const double* jones,
                                double* out) {
    return guarded([&] { field_at_impl(plan, problem, nz, zs, jones, out, "srcwa_plan_conical_field_at", 2); });
}

int srcwa_plan_conical_field_map(srcwa_plan* plan, int problem, int nz, const double* zs, int nx, double y,
                                 const double* jones, double* out) {
    return guarded([&] { field_map_impl(plan, problem, nz, zs, nx, y, jones, out, "srcwa_plan_conical_field_map", 2); });
}

int64_t srcwa_plan_orders(const srcwa_plan* plan) { return plan ? plan->total_orders : -1; }

// The per-instance records hold DIAG_WORDS doubles: the diagnostics report first, then the refinement
// record; each entry point hands out its part.
static void copy_diag_words(srcwa_plan* plan, double* out, int first, int count) {
    hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
    std::vector<double> all(size_t(DIAG_WORDS) * plan->n_inst);
    copy_out(all.data(), plan->diag.p, all.size() * sizeof(double), hipMemcpyDeviceToHost, "D2H", plan);
    for (int i = 0; i < plan->n_inst; ++i)
        for (int k = 0; k < count; ++k) out[size_t(count) * i + k] = all[size_t(DIAG_WORDS) * i + first + k];
}

int srcwa_plan_diagnostics(srcwa_plan* plan, double* out) {
    return guarded([&] {
        if (!plan || !out) invalid("srcwa_plan_diagnostics: null plan or output");
        copy_diag_words(plan, out, 0, DIAG_REPORT_WORDS);
    });
}

int srcwa_plan_refinement(srcwa_plan* plan, double* out) {
    return guarded([&] {
        if (!plan || !out) invalid("srcwa_plan_refinement: null plan or output");
        copy_diag_words(plan, out, DIAG_REPORT_WORDS, DIAG_WORDS - DIAG_REPORT_WORDS);
    });
}

int srcwa_plan_geometry(srcwa_plan* plan, double* out) {
    return guarded([&] {
        if (!plan || !out) invalid("srcwa_plan_geometry: null plan or output");
        if (plan->kind == 1) invalid("srcwa_plan_geometry: planar plans carry no per-order geometry");
        hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(plan->ctx->stream), "hipStreamSynchronize");
        hip_check(hipMemcpyAsync(out, plan->geometry.p, GEO_WORDS * plan->total_orders * sizeof(double),
                                 hipMemcpyDeviceToHost, plan->ctx->stream), "D2H");
        hip_check(hipStreamSynchronize(plan->ctx->stream), "hipStreamSynchronize");
    });
}
int64_t srcwa_plan_device_bytes(const srcwa_plan* plan) { return plan ? plan->bytes() : -1; }
int srcwa_plan_operator_sets(const srcwa_plan* plan) { return plan ? plan->n_ops : -1; }
int srcwa_plan_chunks(const srcwa_plan* plan) { return plan ? static_cast<int>(plan->chunks.size()) : -1; }

int srcwa_plan_stage_ms(const srcwa_plan* plan, double* ms4) {
    return guarded([&] {
        if (!plan || !ms4) invalid("null plan or output");
        if (!plan->timed) invalid("plan has not been executed");
        hip_check(hipEventSynchronize(plan->ev.back()), "hipEventSynchronize");
        double a = 0, b = 0, c = 0;
        if (plan->graph_exec) a = b = c = -1.0;     // a replayed graph carries no stage events
        for (size_t k = 0; k < plan->chunks.size() && !plan->graph_exec; ++k) {
            const hipEvent_t* e = plan->ev.data() + 5 * k;
            float x = 0, y = 0, z = 0;
            hip_check(hipEventElapsedTime(&x, e[0], e[1]), "hipEventElapsedTime");
            hip_check(hipEventElapsedTime(&y, e[1], e[2]), "hipEventElapsedTime");
            hip_check(hipEventElapsedTime(&z, e[3], e[4]), "hipEventElapsedTime");
            a += x;
            b += y;
            c += z;
        }
        float tt = 0;
        hip_check(hipEventElapsedTime(&tt, plan->ev.front(), plan->ev.back()), "hipEventElapsedTime");
        ms4[0] = a;
        ms4[1] = b;
        ms4[2] = c;
        ms4[3] = tt;
    });
}

int srcwa_plan_flops(const srcwa_plan* plan, double* flops2) {
    return guarded([&] {
        if (!plan || !flops2) invalid("null plan or output");
        flops2[0] = plan->flops[0];
        flops2[1] = plan->flops[1];
    });
}

void srcwa_plan_destroy(srcwa_plan* plan) {
    if (!plan) return;
    (void)hipSetDevice(plan->ctx->device);
    (void)hipStreamSynchronize(plan->ctx->stream);
    delete plan;
}

int srcwa_solve1d_batch(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                        int n_problems, const int* stack_index, const double* wavelength, const double* theta_deg,
                        const int* M, const int* pol, double* R, double* T, double* DEr, double* DEt, int* status,
                        int* eig_info) {
    srcwa_plan* plan = nullptr;
    int rc = srcwa_plan1d_create(ctx, n_stacks, stacks, stack_offsets, n_problems, stack_index, wavelength, theta_deg,
                                 M, pol, 0, &plan);
    if (rc != SRCWA_OK) return rc;
    rc = srcwa_plan_execute(plan);
    if (rc == SRCWA_OK) rc = srcwa_plan_results(plan, R, T, DEr, DEt, status, eig_info);
    srcwa_plan_destroy(plan);
    return rc;
}

int srcwa_zgemm_batched(srcwa_context* ctx, int count, const int* m, const int* n, const int* k, const double* A,
                        const double* B, double* C) {
    return dense_batch(ctx, [&](hipStream_t s) {
        if (count < 1 || !m || !n || !k || !A || !B || !C) invalid("srcwa_zgemm_batched: empty or null inputs");
        std::vector<DenseItem> items(count);
        int64_t off = 0, ao = 0, bo = 0, co = 0;
        for (int i = 0; i < count; ++i) {
            if (m[i] < 0 || n[i] < 0 || k[i] < 0) invalid("srcwa_zgemm_batched: negative size");
            DenseItem& it = items[i];
            it.n = m[i];
            it.nrhs = n[i];
            it.k = k[i];
            it.a = off;
            off += align16(int64_t(m[i]) * k[i]);
            it.b = off;
            off += align16(int64_t(k[i]) * n[i]);
            it.c = off;
            off += align16(int64_t(m[i]) * n[i]);
        }
        DevBuf<cd> arena;
        arena.alloc(off);
        DevBuf<DenseItem> d_items;
        d_items.upload(items, s);
        for (int i = 0; i < count; ++i) {
            const int64_t as = int64_t(m[i]) * k[i], bs = int64_t(k[i]) * n[i];
            hip_check(hipMemcpyAsync(arena.p + items[i].a, A + 2 * ao, as * 16, hipMemcpyHostToDevice, s), "H2D");
            hip_check(hipMemcpyAsync(arena.p + items[i].b, B + 2 * bo, bs * 16, hipMemcpyHostToDevice, s), "H2D");
            ao += as;
            bo += bs;
        }
        launch_zgemm_batched(s, count, d_items.p, arena.p);
        hip_check(hipGetLastError(), "kernel launch");
        for (int i = 0; i < count; ++i) {
            const int64_t cs = int64_t(m[i]) * n[i];
            hip_check(hipMemcpyAsync(C + 2 * co, arena.p + items[i].c, cs * 16, hipMemcpyDeviceToHost, s), "D2H");
            co += cs;
        }
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}

namespace {
// refine == nullptr: the FP64 LU (srcwa_zgesv_batched); else FP32 LU + FP64 refinement, 3 doubles per item
int zgesv_batched(srcwa_context* ctx, int count, const int* n, const int* nrhs, const double* A, const double* B,
                  double* X, int* info, double* refine) {
    const char* name = refine ? "srcwa_zgesv_batched_mixed" : "srcwa_zgesv_batched";
    return dense_batch(ctx, [&](hipStream_t s) {
        if (count < 1 || !n || !nrhs || !A || !B || !X || !info)
            invalid(std::string(name) + ": empty or null inputs");
        std::vector<DenseItem> items(count);
        int64_t off = 0;
        for (int i = 0; i < count; ++i) {
            if (n[i] < 1 || nrhs[i] < 0) invalid(std::string(name) + ": bad size");
            DenseItem& it = items[i];
            it.n = n[i];
            it.nrhs = nrhs[i];
            it.k = 0;
            it.a = off;
            off += align16(int64_t(n[i]) * n[i]);
            it.b = off;
            off += align16(int64_t(n[i]) * nrhs[i]);
            it.scratch = off;
            off += align16(n[i]);
            if (refine) off += align16(int64_t(n[i]) * n[i]) + refine_words(n[i], nrhs[i]);
        }
        DevBuf<cd> arena;
        arena.alloc(off);
        DevBuf<DenseItem> d_items;
        d_items.upload(items, s);
        DevBuf<int> d_info;
        d_info.alloc(count);
        int64_t ao = 0, bo = 0;
        for (int i = 0; i < count; ++i) {
            const int64_t as = int64_t(n[i]) * n[i], bs = int64_t(n[i]) * nrhs[i];
            hip_check(hipMemcpyAsync(arena.p + items[i].a, A + 2 * ao, as * 16, hipMemcpyHostToDevice, s), "H2D");
            hip_check(hipMemcpyAsync(arena.p + items[i].b, B + 2 * bo, bs * 16, hipMemcpyHostToDevice, s), "H2D");
            ao += as;
            bo += bs;
        }
        DevBuf<double> d_refine;
        if (refine) {
#if SRCWA_MIXED_PRECISION
            d_refine.alloc(size_t(3) * count);
            launch_zgesv_mixed_batched(s, count, d_items.p, arena.p, d_info.p, d_refine.p);
#endif
        } else {
            launch_zgesv_batched(s, count, d_items.p, arena.p, d_info.p);
        }
        hip_check(hipGetLastError(), "kernel launch");
        if (refine)
            hip_check(hipMemcpyAsync(refine, d_refine.p, size_t(3) * count * sizeof(double), hipMemcpyDeviceToHost, s),
                      "D2H");
        bo = 0;
        for (int i = 0; i < count; ++i) {
            const int64_t bs = int64_t(n[i]) * nrhs[i];
            hip_check(hipMemcpyAsync(X + 2 * bo, arena.p + items[i].b, bs * 16, hipMemcpyDeviceToHost, s), "D2H");
            bo += bs;
        }
        hip_check(hipMemcpyAsync(info, d_info.p, count * sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}
}  // namespace

int srcwa_zgesv_batched(srcwa_context* ctx, int count, const int* n, const int* nrhs, const double* A, const double* B,
                        double* X, int* info) {
    return zgesv_batched(ctx, count, n, nrhs, A, B, X, info, nullptr);
}

int srcwa_zgesv_batched_mixed(srcwa_context* ctx, int count, const int* n, const int* nrhs, const double* A,
                              const double* B, double* X, int* info, double* refine) {
    if (!SRCWA_MIXED_PRECISION) return guarded([&] { throw Failure{SRCWA_EUNSUPPORTED, kMixedOff}; });
    if (!refine) return guarded([&] { invalid("srcwa_zgesv_batched_mixed: null refinement output"); });
    return zgesv_batched(ctx, count, n, nrhs, A, B, X, info, refine);
}

int srcwa_zgeev_batched(srcwa_context* ctx, int count, const int* n, const double* A, double* w, double* V, int* info) {
    return srcwa_zgeev_batched_iter(ctx, count, n, A, w, V, info, 30);
}

int srcwa_zgeev_batched_iter(srcwa_context* ctx, int count, const int* n, const double* A, double* w, double* V,
                             int* info, int max_iter) {
    return dense_batch(ctx, [&](hipStream_t s) {
        if (count < 1 || !n || !A || !w || !V || !info) invalid("srcwa_zgeev_batched: empty or null inputs");
        if (max_iter < 1 || max_iter > 100000) invalid("srcwa_zgeev_batched: max_iter must lie in 1..100000");
        std::vector<DenseItem> items(count);
        int64_t off = 0;
        for (int i = 0; i < count; ++i) {
            if (n[i] < 1 || n[i] > EIG_MAX_N_HOST) invalid("srcwa_zgeev_batched: order out of range");
            DenseItem& it = items[i];
            const int64_t nn = int64_t(n[i]) * n[i];
            it.n = n[i];
            it.nrhs = 0;
            it.k = 0;
            it.a = off;
            off += align16(nn);
            it.w = off;
            off += align16(n[i]);
            it.scratch = off;
            off += align16(2 * nn + 2 * n[i]);
        }
        DevBuf<cd> arena;
        arena.alloc(off);
        DevBuf<DenseItem> d_items;
        d_items.upload(items, s);
        DevBuf<int> d_info;
        d_info.alloc(count);
        int64_t ao = 0;
        for (int i = 0; i < count; ++i) {
            const int64_t as = int64_t(n[i]) * n[i];
            hip_check(hipMemcpyAsync(arena.p + items[i].a, A + 2 * ao, as * 16, hipMemcpyHostToDevice, s), "H2D");
            ao += as;
        }
        launch_zgeev_batched(s, count, d_items.p, arena.p, d_info.p, max_iter, *std::max_element(n, n + count));
        hip_check(hipGetLastError(), "kernel launch");
        ao = 0;
        int64_t wo = 0;
        for (int i = 0; i < count; ++i) {
            const int64_t as = int64_t(n[i]) * n[i];
            hip_check(hipMemcpyAsync(V + 2 * ao, arena.p + items[i].a, as * 16, hipMemcpyDeviceToHost, s), "D2H");
            hip_check(hipMemcpyAsync(w + 2 * wo, arena.p + items[i].w, n[i] * 16, hipMemcpyDeviceToHost, s), "D2H");
            ao += as;
            wo += n[i];
        }
        hip_check(hipMemcpyAsync(info, d_info.p, count * sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}

}  // extern "C"

// ---- crossed gratings (rcwa2d.hpp:953-1026, rcwa2d_operators.hpp:244-320) ------------------------
namespace {

void check_index_2d(double re, double im) {
    if (!finite_c(re, im) || std::hypot(re, im) == 0.0 || re < 0.0 || im < 0.0)
        invalid("rcwa2d: indices must be finite, nonzero, with Re n >= 0 and Im n >= 0");
}

void check_edge_list(const double* e, int count, double period, const char* what) {
    if (count < 2) invalid(std::string("rcwa2d: ") + what + " need at least two edges");
    for (int j = 0; j < count; ++j)
        if (!std::isfinite(e[j]) || (j > 0 && e[j] <= e[j - 1]))
            invalid(std::string("rcwa2d: ") + what + " must be finite and strictly increasing");
    if (std::abs((e[count - 1] - e[0]) - period) > 1e-9 * period)
        invalid(std::string("rcwa2d: ") + what + " must span exactly one period");
}

double floor_mod_host(double a, double b) {
    double m = std::fmod(a, b);
    if (m != 0.0) {
        if ((b < 0.0) != (m < 0.0)) m += b;
    } else {
        m = std::copysign(0.0, b);
    }
    return m;
}

struct Layer2DInfo {
    bool patterned;
    int K = 0, maxkx = 1;
    std::vector<double> cuts;     // merged x cuts (patterned layers)
    int64_t rec = 0, rec_words = 0;   // the layer's words after its thickness in the stack record
};

// Parses and validates one crossed-grating stack record; returns per-layer info and the record length.
std::vector<Layer2DInfo> parse_stack2d(const double* p, int64_t avail, int64_t& words) {
    if (avail < 6) invalid("rcwa2d: malformed stack record");
    const double Lx = p[0], Ly = p[1];
    if (!(std::isfinite(Lx) && Lx > 0.0) || !(std::isfinite(Ly) && Ly > 0.0))
        invalid("rcwa2d: wavelength and periods must be finite and positive");
    if (!std::isfinite(p[2]) || p[2] <= 0.0) invalid("rcwa2d: the incident medium must have a real positive index");
    check_index_2d(p[3], p[4]);
    const int L = static_cast<int>(p[5]);
    if (!(p[5] >= 0.0) || L > 128) invalid("rcwa2d: a stack holds 0..128 layers");
    std::vector<Layer2DInfo> out;
    int64_t k = 6;
    auto need = [&](int64_t n) {
        if (k + n > avail) invalid("rcwa2d: malformed stack record");
    };
    for (int j = 0; j < L; ++j) {
        need(2);
        const double d = p[k];
        const int kind = static_cast<int>(p[k + 1]);
        const int64_t start = k + 1;
        k += 2;
        if (!std::isfinite(d) || d < 0.0) invalid("rcwa2d: layer thickness must be finite and >= 0");
        Layer2DInfo info;
        if (kind == 0) {
            need(2);
            check_index_2d(p[k], p[k + 1]);
            k += 2;
            info.patterned = false;
        } else if (kind == 1) {
            need(1);
            const int K = static_cast<int>(p[k]);
            if (K < 1 || K > 256) invalid("rcwa2d: a patterned layer needs 1..256 y segments");
            need(1 + K + 1);
            check_edge_list(p + k + 1, K + 1, Ly, "y-edges");
            k += 1 + K + 1;
            info.patterned = true;
            info.K = K;
            std::vector<double> cuts;
            for (int y = 0; y < K; ++y) {
                need(1);
                const int Kx = static_cast<int>(p[k]);
                if (Kx < 1 || Kx > 4096) invalid("rcwa2d: an x-profile needs K indices and K+1 edges");
                need(1 + (Kx + 1) + 2 * Kx);
                const double* xe = p + k + 1;
                const double* v = xe + Kx + 1;
                check_edge_list(xe, Kx + 1, Lx, "x-edges");
                for (int i = 0; i < Kx; ++i) {
                    check_index_2d(v[2 * i], v[2 * i + 1]);
                    const std::complex<double> n{v[2 * i], v[2 * i + 1]};
                    const std::complex<double> eps = n * n;
                    if (!finite_c(eps.real(), eps.imag()) || eps == std::complex<double>{})
                        invalid("rcwa2d: eps = n^2 overflows or vanishes in a patterned layer");
                    const std::complex<double> inv = 1.0 / eps;
                    if (!finite_c(inv.real(), inv.imag())) invalid("rcwa2d: 1/eps overflows in a patterned layer");
                }
                for (int i = 0; i <= Kx; ++i) cuts.push_back(floor_mod_host(xe[i], Lx));
                info.maxkx = std::max(info.maxkx, Kx);
                k += 1 + (Kx + 1) + 2 * Kx;
            }
            // merge_x_cuts (rcwa2d_operators.hpp:160-174)
            cuts.push_back(0.0);
            cuts.push_back(Lx);
            std::sort(cuts.begin(), cuts.end());
            cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
            for (double& c : cuts) c = std::rint(c * 1e12) / 1e12;
            std::sort(cuts.begin(), cuts.end());
            cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
            cuts.erase(std::remove_if(cuts.begin(), cuts.end(), [&](double c) { return !(c >= -1e-12 && c <= Lx + 1e-12); }),
                       cuts.end());
            if (cuts.empty() || cuts.front() > 1e-12) cuts.insert(cuts.begin(), 0.0);
            if (cuts.back() < Lx - 1e-12) cuts.push_back(Lx);
            info.cuts = std::move(cuts);
            info.rec = start;
            info.rec_words = k - start;
        } else {
            invalid("rcwa2d: unknown layer kind");
        }
        out.push_back(std::move(info));
    }
    words = k;
    return out;
}

// cd words of k_ops2d's scratch (same layout as the kernel)
int64_t ops2d_scratch_words(int K, int Kc, int maxkx, int M1, int M2) {
    const int64_t n1 = 2 * M1 + 1, n2 = 2 * M2 + 1, n1s = n1 * n1, n2s = n2 * n2;
    int64_t w = int64_t(K) * (4 * M2 + 1) + 2 * int64_t(K) * n1s + int64_t(Kc) * (4 * M1 + 1) + int64_t(Kc) * n2s;
    w += std::max(int64_t(K) * n1s, int64_t(Kc) * n2s);
    w += std::max(int64_t(4 * M2 + 1) * n1s, int64_t(4 * M1 + 1) * n2s);
    const int64_t coefsz = std::max<int64_t>(maxkx + 4 * M1 + 1, K + 4 * M2 + 1);
    w += coefsz * std::max(K, Kc);
    w += (std::max(int64_t(K) * n1, int64_t(Kc) * n2) + 3) / 4 + 16;
    return w;
}

enum RKind { RZ, RB, RD };

// Workspace of the crossed-grating recursion and its flops, along the structural kinds.
PlanarCost crossed_cost(int N, const std::vector<bool>& patterned) {
    const int L = static_cast<int>(patterned.size()) - 2;
    const double n = 2.0 * N, n3 = n * n * n, h3 = double(N) * N * N;
    const int64_t nn = 4 * int64_t(N) * N, b4 = 4 * int64_t(N);
    PlanarCost c{0, 0.0, 0.0};
    int64_t persist = 2 * align16(nn);
    int r = RZ;
    const double LU = 8.0 / 3.0 * n3, SOLVE_N = 8.0 * n3, GEMM = 8.0 * n3;
    for (int b = 2 * L; b >= 1; --b) {
        const int j = b / 2;
        if (b % 2 == 0) {
            const bool homog = !patterned[j] && !patterned[j + 1];
            if (homog) {
                if (r != RD) {
                    persist += align16(b4);
                    r = RB;
                } else {
                    persist += 2 * align16(b4) + align16(nn);
                    c.flops_other += LU + SOLVE_N;
                    r = RD;
                }
            } else {
                persist += 3 * align16(nn);
                if (patterned[j + 1]) c.flops_other += 2 * (LU + SOLVE_N);
                if (r == RD) c.flops_other += 2 * GEMM;
                c.flops_other += LU + SOLVE_N;
                r = RD;
            }
        } else {
            persist += align16(b4);
            if (r != RD) r = RB;
        }
    }
    if (patterned[1]) c.flops_other += LU + 8.0 * n * n + (r == RD ? 2 * GEMM : 0.0);
    for (int j = 1; j <= L; ++j)
        if (patterned[j]) {
            c.flops_eig += 100.0 * n3;
            c.flops_other += 8.0 / 3.0 * h3 + 8.0 * h3 + 32.0 * h3 + 32.0 * h3;   // Ei, H, V
        }
    c.bump_words = persist + 10 * nn + 64 * int64_t(N) + 4096;
    return c;
}

}  // namespace

static int plan2d_create_impl(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                                   int n_problems, const int* stack_index, const double* wavelength,
                                   const double* theta_deg, const double* phi_deg, const int* M1, const int* M2,
                                   const double* jones, const int* factorization, int64_t max_workspace_bytes,
                                   srcwa_plan** out, int64_t* need) {
    return guarded([&] {
        if (!ctx || !out) invalid("srcwa_plan2d_create: null context or output");
        if (n_stacks < 1 || n_problems < 1 || !stacks || !stack_offsets || !stack_index || !wavelength ||
            !theta_deg || !phi_deg || !M1 || !M2 || !jones)
            invalid("srcwa_plan2d_create: empty or null inputs");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        std::vector<int64_t> sorted(stack_offsets, stack_offsets + n_stacks);
        for (int64_t o : sorted)
            if (o < 0) invalid("rcwa2d: malformed stack offsets");
        std::sort(sorted.begin(), sorted.end());
        std::vector<std::vector<Layer2DInfo>> infos(n_stacks);
        int64_t stack_words = 0;
        for (int s = 0; s < n_stacks; ++s) {
            const int64_t off = stack_offsets[s];
            auto it = std::upper_bound(sorted.begin(), sorted.end(), off);
            const int64_t avail = it == sorted.end() ? (int64_t(1) << 40) : *it - off;
            int64_t words = 0;
            infos[s] = parse_stack2d(stacks + off, avail, words);
            stack_words = std::max(stack_words, off + words);
        }
        // stack records, then every patterned layer's x-cut record [Kc, cuts]
        std::vector<double> h_stacks(stacks, stacks + stack_words);
        std::vector<std::vector<int64_t>> cut_off(n_stacks);
        for (int s = 0; s < n_stacks; ++s) {
            cut_off[s].assign(infos[s].size(), -1);
            for (size_t j = 0; j < infos[s].size(); ++j)
                if (infos[s][j].patterned) {
                    cut_off[s][j] = static_cast<int64_t>(h_stacks.size());
                    h_stacks.push_back(static_cast<double>(infos[s][j].cuts.size() - 1));
                    for (double c : infos[s][j].cuts) h_stacks.push_back(c);
                }
        }
        auto plan = std::unique_ptr<srcwa_plan>(new srcwa_plan);
        plan->bind_pool(ctx->pool);
        plan->ctx = ctx;
        plan->kind = 3;
        plan->n_inst = n_problems;
        plan->flags = ctx->plan_flags;
        if (plan->flags & PLAN_FIELDS)
            throw Failure{SRCWA_EUNSUPPORTED, "rcwa: near fields are captured for planar and conical plans only in this build"};
        const bool diag = plan->flags & PLAN_DIAGNOSTICS;
        const int64_t budget = max_workspace_bytes > 0 ? max_workspace_bytes : (int64_t(64) << 30);
        const int64_t budget_words = budget / int64_t(sizeof(cd));
        std::vector<Inst2D> h_insts(n_problems);
        std::vector<LayerTask> h_tasks;
        std::vector<int64_t> h_layouts;
        int64_t orders = 0, chunk_ws = 0, max_ws = 0;
        PlanChunk cur{0, 0, 0, 0};
        OpSharing sharing;
        sharing.enabled = !(plan->flags & (PLAN_DIAGNOSTICS | PLAN_POINT_OPS));
        const int64_t limit = (std::numeric_limits<int>::max() - 1) / 4;
        for (int i = 0; i < n_problems; ++i) {
            const int s = stack_index[i];
            if (s < 0 || s >= n_stacks) invalid("rcwa2d: stack index out of range");
            if (M1[i] < 0 || M2[i] < 0 || (2LL * M1[i] + 1) * (2LL * M2[i] + 1) > limit)
                invalid("rcwa2d: invalid truncation order M1, M2");
            if (!std::isfinite(wavelength[i]) || wavelength[i] <= 0.0)
                invalid("rcwa2d: wavelength and periods must be finite and positive");
            if (!std::isfinite(theta_deg[i]) || std::abs(theta_deg[i]) >= 90.0)
                invalid("rcwa2d: incidence angle must lie strictly between -90 and 90 degrees");
            if (!std::isfinite(phi_deg[i])) invalid("rcwa2d: azimuth angle must be finite");
            for (int q = 0; q < 4; ++q)
                if (!std::isfinite(jones[4 * i + q])) invalid("rcwa2d: the Jones vector must be finite");
            const int fz = factorization ? factorization[i] : 0;
            if (fz != 0 && fz != 1) invalid("rcwa2d: unknown factorization");
            const std::vector<Layer2DInfo>& li = infos[s];
            const int L = static_cast<int>(li.size());
            const int N = (2 * M1[i] + 1) * (2 * M2[i] + 1);
            if (2 * N > EIG_MAX_N_HOST)
                throw Failure{SRCWA_EUNSUPPORTED, "rcwa2d: harmonic count above this build's limit"};
            std::vector<bool> patterned(L + 2, false);
            for (int j = 0; j < L; ++j) patterned[j + 1] = li[j].patterned;
            const PlanarCost cost = crossed_cost(N, patterned);
            plan->flops[0] += cost.flops_eig;
            plan->flops[1] += cost.flops_other;
            const int64_t NN = int64_t(N) * N;
            int64_t w = 0;
            std::vector<int64_t> lay(x2_layout_size(L), -1);
            lay[0] = w;
            w += align16(4 * N + 1);
            for (int r = 0; r < L + 2; ++r) {
                lay[x2_q(r)] = w;
                w += align16(2 * N);
                if (!patterned[r]) {
                    lay[x2_v(r)] = w;
                    w += align16(4 * N);
                }
            }
            for (int j = 0; j < L; ++j)
                if (li[j].patterned) {
                    const int Kc = static_cast<int>(li[j].cuts.size()) - 1;
                    const int64_t big = align16(std::max<int64_t>(12 * NN, ops2d_scratch_words(li[j].K, Kc, li[j].maxkx,
                                                                                              M1[i], M2[i])));
                    lay[x2_slot(L, j)] = w;
                    lay[x2_big(L, j)] = big;
                    lay[x2_cuts(L, j)] = cut_off[s][j];
                    w += align16(3 * NN + big + 7 * int64_t(N));
                    if (diag) w += align16(4 * NN);      // a copy of H for the eigen residual
                }
            const int64_t ws_lo = w;
            w += cost.bump_words + ((plan->flags & PLAN_MIXED) ? mixed_solve_words(2 * N, 2 * N) : 0);
            if (i > cur.i0 && chunk_ws + w > budget_words) {
                cur.i1 = i;
                cur.t1 = static_cast<int>(h_tasks.size());
                plan->chunks.push_back(cur);
                max_ws = std::max(max_ws, chunk_ws);
                cur = PlanChunk{i, i, cur.t1, cur.t1};
                chunk_ws = 0;
            }
            const int first_task = static_cast<int>(h_tasks.size());
            for (int j = 0; j < L; ++j)
                if (li[j].patterned) {
                    h_tasks.push_back(LayerTask{i - cur.i0, j});
                    // Eps, Cx, Cy (rcwa2d_operators.hpp:349-429) and Eps^-1 (rcwa2d.hpp:412-420)
                    const double* st = stacks + stack_offsets[s];
                    std::vector<uint64_t> key;
                    for (double v : {st[0], st[1], double(M1[i]), double(M2[i]), double(fz)}) OpSharing::push(key, v);
                    for (int64_t k = 0; k < li[j].rec_words; ++k) OpSharing::push(key, st[li[j].rec + k]);
                    const int Kc = static_cast<int>(li[j].cuts.size()) - 1;
                    sharing.add(key, i, j,
                                crossed_ops_words(N, ops2d_scratch_words(li[j].K, Kc, li[j].maxkx, M1[i], M2[i])));
                }
            const int64_t layout_off = static_cast<int64_t>(h_layouts.size());
            for (int r = 0; r < x2_layout_size(L); ++r) {
                int64_t v = lay[r];
                const bool is_offset = r < x2_cuts(L, 0) || L == 0;
                if (is_offset && v >= 0) v += chunk_ws;
                h_layouts.push_back(v);
            }
            Inst2D& in = h_insts[i];
            in.M1 = M1[i];
            in.M2 = M2[i];
            in.N = N;
            plan->max_n = std::max(plan->max_n, N);
            plan->min_n = std::min(plan->min_n, N);
            in.L = L;
            in.first_task = first_task - cur.t0;
            in.n_tasks = static_cast<int>(h_tasks.size()) - first_task;
            in.laurent = fz;
            in.flags = plan->flags;
            in.wavelength = wavelength[i];
            in.theta_deg = theta_deg[i];
            in.phi_deg = phi_deg[i];
            for (int q = 0; q < 4; ++q) in.jones[q] = jones[4 * i + q];
            in.stack_off = stack_offsets[s];
            in.layout_off = layout_off;
            in.out_off = orders;
            in.ws_lo = chunk_ws + ws_lo;
            in.ws_hi = chunk_ws + w;
            chunk_ws += w;
            orders += N;
        }
        cur.i1 = n_problems;
        cur.t1 = static_cast<int>(h_tasks.size());
        plan->chunks.push_back(cur);
        max_ws = std::max(max_ws, chunk_ws);
        plan->n_tasks = static_cast<int>(h_tasks.size());
        plan->total_orders = orders;
        std::vector<LayerOps> h_ops;
        const int64_t op_words = sharing.assign(h_tasks, h_ops);
        if (need) {
            const auto b = [](size_t x) { return static_cast<int64_t>(srcwa_pool::round(x)); };
            *need = b(h_insts.size() * sizeof(Inst2D)) + b(h_tasks.size() * sizeof(LayerTask)) +
                    b(std::max<size_t>(1, h_tasks.size()) * sizeof(int)) +
                    b(h_stacks.size() * sizeof(double)) + b(h_layouts.size() * sizeof(int64_t)) +
                    b((plan->chunks.size() > 1 ? 2 : 1) * max_ws * sizeof(cd)) + b(4 * std::max(1, plan->n_tasks) * sizeof(int)) +
                    2 * b(8 * orders * sizeof(double)) + 2 * b(orders * sizeof(double)) +
                    b(GEO_WORDS * orders * sizeof(double)) + 2 * b(n_problems * sizeof(int)) +
                    diag_bytes(n_problems, plan->n_tasks) + ops_bytes(h_ops, op_words);
            return;
        }
        hipStream_t st = ctx->stream;
        plan->insts2d.upload(h_insts, st);
        plan->tasks.upload(h_tasks, st);
        plan->upload_ops(h_ops, op_words, st);
        plan->eig_order.upload(longest_first(plan->chunks, h_tasks, h_insts), st);
        plan->stacks.upload(h_stacks, st);
        plan->layouts.upload(h_layouts, st);
        plan->region_words = max_ws;
        plan->ws.alloc(static_cast<size_t>(plan->chunks.size() > 1 ? 2 * max_ws : max_ws));
        plan->task_info.alloc(4 * std::max(1, plan->n_tasks));
        plan->R.alloc(8 * orders);
        plan->T.alloc(8 * orders);
        plan->geometry.alloc(GEO_WORDS * orders);
        plan->DEr.alloc(orders);
        plan->DEt.alloc(orders);
        plan->status.alloc(n_problems);
        plan->eig_info.alloc(n_problems);
        plan->alloc_diag();
        plan->make_events();
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        *out = plan.release();
    });
}

extern "C" int srcwa_plan2d_create(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                                   int n_problems, const int* stack_index, const double* wavelength,
                                   const double* theta_deg, const double* phi_deg, const int* M1, const int* M2,
                                   const double* jones, const int* factorization, int64_t max_workspace_bytes,
                                   srcwa_plan** out) {
    return plan2d_create_impl(ctx, n_stacks, stacks, stack_offsets, n_problems, stack_index, wavelength, theta_deg,
           phi_deg, M1, M2, jones, factorization, max_workspace_bytes, out, nullptr);
}

extern "C" int srcwa_plan2d_bytes(srcwa_context* ctx, int n_stacks, const double* stacks, const int64_t* stack_offsets,
                                   int n_problems, const int* stack_index, const double* wavelength,
                                   const double* theta_deg, const double* phi_deg, const int* M1, const int* M2,
                                   const double* jones, const int* factorization, int64_t max_workspace_bytes,
                                   int64_t* bytes) {
    if (!bytes) {
        g_error = "null output";
        return SRCWA_EINVAL;
    }
    srcwa_plan* unused = nullptr;
    return plan2d_create_impl(ctx, n_stacks, stacks, stack_offsets, n_problems, stack_index, wavelength, theta_deg,
           phi_deg, M1, M2, jones, factorization, max_workspace_bytes, &unused, bytes);
}
