// Host emulation of one HIP workgroup with real threads, for checking the 1-D
// kernels' barrier discipline on the CPU (built with ThreadSanitizer by
// tests/test_emu_mt.py).  Each emulated GPU thread is a host thread with its
// own threadIdx; __syncthreads is a barrier over the workgroup; LDS
// (__shared__ and the dynamic array) is memory shared by all threads;
// __shfl_xor exchanges values inside a 64-thread wave through a buffer and a
// per-wave barrier; atomics are host atomics.  A thread that returns from the
// kernel leaves every barrier, as an exited wave does on the GPU.  Any two
// threads touching the same memory between barriers without atomics is a
// data race that ThreadSanitizer reports.  Test infrastructure only.
#pragma once

#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#define __device__
#define __host__
#define __global__
#define __shared__ static
#define __forceinline__ inline
#define __align__(n) __attribute__((aligned(n)))
#define __launch_bounds__(...)
#define AWCM_DYNAMIC_LDS extern

struct EmuDim3 {
    unsigned x = 0, y = 0, z = 0;
};
inline thread_local EmuDim3 threadIdx;
inline EmuDim3 blockIdx;
inline EmuDim3 blockDim{1, 1, 1};
inline EmuDim3 gridDim{1, 1, 1};

typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;

namespace emu_mt {
inline std::barrier<>* block_bar = nullptr;
inline std::vector<std::unique_ptr<std::barrier<>>> wave_bars;
inline std::vector<double> xbuf;  // one shuffle slot per thread
}  // namespace emu_mt

inline void __syncthreads() { emu_mt::block_bar->arrive_and_wait(); }

inline double __shfl_xor(double v, int o, int = 64) {
    const unsigned t = threadIdx.x, w = t / 64;
    std::barrier<>& b = *emu_mt::wave_bars[w];
    emu_mt::xbuf[t] = v;
    b.arrive_and_wait();
    const unsigned src = w * 64 + ((t % 64) ^ unsigned(o));
    const double r = src < blockDim.x ? emu_mt::xbuf[src] : v;
    b.arrive_and_wait();
    return r;
}

inline long long __double_as_longlong(double v) {
    long long r;
    std::memcpy(&r, &v, 8);
    return r;
}
template <class T, class U>
inline T atomicAdd(T* p, U v) {
    return __atomic_fetch_add(p, T(v), __ATOMIC_SEQ_CST);
}
template <class T, class U>
inline T atomicOr(T* p, U v) {
    return __atomic_fetch_or(p, T(v), __ATOMIC_SEQ_CST);
}
template <class T, class U>
inline T atomicMax(T* p, U v) {
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old < T(v) && !__atomic_compare_exchange_n(p, &old, T(v), false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
using std::max;
using std::min;

// Runs f() as one workgroup of n threads (n a multiple of 64).
template <class F>
void emu_launch(unsigned n, F f) {
    blockDim.x = n;
    std::barrier<> bb(n);
    emu_mt::block_bar = &bb;
    emu_mt::wave_bars.clear();
    for (unsigned w = 0; w * 64 < n; w++) emu_mt::wave_bars.emplace_back(new std::barrier<>(std::min(64u, n - 64 * w)));
    emu_mt::xbuf.assign(n, 0.0);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < n; t++)
        th.emplace_back([&, t] {
            threadIdx.x = t;
            f();
            emu_mt::wave_bars[t / 64]->arrive_and_drop();
            bb.arrive_and_drop();
        });
    for (auto& x : th) x.join();
    emu_mt::block_bar = nullptr;
}
