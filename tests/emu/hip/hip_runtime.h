// CPU emulation of the HIP device subset that stablercwa_amd/csrc/*.hpp use, for tests only: one workgroup
// of 256 std::threads, __syncthreads as a barrier of all 256, wave collectives (shuffles, readlane, the
// 16x16x4 MFMAs) through per-wave exchange buffers and a barrier of the wave's 64 threads, __shared__
// variables as statics (one workgroup runs at a time).  readfirstlane is the identity: the code applies it
// to wave-uniform values only, also in divergent code where a collective would not be reached by all lanes.
#pragma once
#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(x) __attribute__((aligned(x)))

using std::max;
using std::min;
using std::isfinite;
using std::isnan;
using std::isinf;

struct EmuDim3 {
    unsigned x = 0, y = 0, z = 0;
};
extern thread_local EmuDim3 threadIdx;
extern EmuDim3 blockIdx;

namespace emu {
constexpr int NT = 256;
extern std::barrier<>* block_bar;
extern std::barrier<>* wave_bar[4];
extern double xbuf_d[4][64];
extern long long xbuf_l[4][64];
extern double mf_a[4][64], mf_b[4][64];
extern double mf_c[4][64][4];
inline int lane() { return threadIdx.x & 63; }
inline int wave() { return threadIdx.x >> 6; }
inline void wsync() { wave_bar[wave()]->arrive_and_wait(); }
}  // namespace emu

inline void __syncthreads() { emu::block_bar->arrive_and_wait(); }

inline long long emu_xchg_l(long long v, int src_lane) {
    const int w = emu::wave();
    emu::xbuf_l[w][emu::lane()] = v;
    emu::wsync();
    const long long r = emu::xbuf_l[w][src_lane & 63];
    emu::wsync();
    return r;
}
inline double __shfl_xor(double v, int mask, int width = 64) {
    long long b;
    std::memcpy(&b, &v, 8);
    b = emu_xchg_l(b, emu::lane() ^ mask);
    double r;
    std::memcpy(&r, &b, 8);
    return r;
}
inline int __shfl_xor(int v, int mask, int width = 64) { return static_cast<int>(emu_xchg_l(v, emu::lane() ^ mask)); }
inline float __shfl_xor(float v, int mask, int width = 64) { return static_cast<float>(__shfl_xor(static_cast<double>(v), mask, width)); }
inline int emu_readlane(int v, int l) { return static_cast<int>(emu_xchg_l(v, l)); }
inline int emu_readfirstlane(int v) { return v; }   // used on wave-uniform values, also in divergent code
inline void emu_wave_barrier() { emu::wsync(); }
#define __builtin_amdgcn_readlane emu_readlane
#define __builtin_amdgcn_readfirstlane emu_readfirstlane
#define __builtin_amdgcn_wave_barrier emu_wave_barrier
#define __builtin_amdgcn_fence(a, b) std::atomic_thread_fence(std::memory_order_seq_cst)
#define __builtin_amdgcn_s_sleep(x) ((void)0)

inline long long __double_as_longlong(double v) {
    long long b;
    std::memcpy(&b, &v, 8);
    return b;
}
inline double __longlong_as_double(long long b) {
    double v;
    std::memcpy(&v, &b, 8);
    return v;
}
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }

typedef double emu_d4 __attribute__((ext_vector_type(4)));
typedef float emu_f4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA: A[i][k] from lane (k*16 + i), B[k][j] from lane (k*16 + j); output element r of lane l at
// row l/16 + 4 r (f64) or 4 (l/16) + r (f32) of column l % 16.  EMU_F32_LAYOUT_F64 = true gives the f32
// instruction the f64 layout instead (a deliberately wrong layout, to exercise the refinement's fallback).
#ifndef EMU_F32_LAYOUT_F64
#define EMU_F32_LAYOUT_F64 false
#endif
template <class V, class T>
inline V emu_mfma(T a, T b, V c, bool f64_layout) {
    const int w = emu::wave(), l = emu::lane();
    emu::mf_a[w][l] = a;
    emu::mf_b[w][l] = b;
    for (int r = 0; r < 4; ++r) emu::mf_c[w][l][r] = c[r];
    emu::wsync();
    V d;
    const int j = l % 16;
    for (int r = 0; r < 4; ++r) {
        const int i = f64_layout ? (l / 16 + 4 * r) : (4 * (l / 16) + r);
        double s = c[r];
        for (int k = 0; k < 4; ++k) s += emu::mf_a[w][k * 16 + i] * emu::mf_b[w][k * 16 + j];
        d[r] = static_cast<T>(s);
    }
    emu::wsync();
    return d;
}
inline emu_d4 emu_mfma_f64(double a, double b, emu_d4 c, int, int, int) { return emu_mfma<emu_d4, double>(a, b, c, true); }
inline emu_f4 emu_mfma_f32(float a, float b, emu_f4 c, int, int, int) { return emu_mfma<emu_f4, float>(a, b, c, EMU_F32_LAYOUT_F64); }
#define __builtin_amdgcn_mfma_f64_16x16x4f64 emu_mfma_f64
#define __builtin_amdgcn_mfma_f32_16x16x4f32 emu_mfma_f32

typedef int hipError_t;
typedef void* hipStream_t;
