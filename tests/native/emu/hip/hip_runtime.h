// Host emulation of the few HIP device features the 1-D kernels use, for
// running those kernels on the CPU in tests: one "thread" (threadIdx.x = 0,
// blockDim.x = 1), barriers as no-ops, LDS as ordinary memory.  Valid for
// kernels whose passes are race-free between barriers, which is how the 1-D
// solver and adjoint kernels are written.  Test infrastructure only.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __shared__
#define __forceinline__ inline
#define __align__(n) __attribute__((aligned(n)))
#define __launch_bounds__(...)

struct EmuDim3 {
    unsigned x = 0, y = 0, z = 0;
};
inline EmuDim3 threadIdx;
inline EmuDim3 blockIdx;
inline EmuDim3 blockDim{1, 1, 1};
inline EmuDim3 gridDim{1, 1, 1};

typedef void* hipStream_t;

inline void __syncthreads() {}
inline double __shfl_xor(double v, int, int = 64) { return v; }
inline long long __double_as_longlong(double v) {
    long long r;
    std::memcpy(&r, &v, 8);
    return r;
}
template <class T, class U>
inline T atomicAdd(T* p, U v) {
    const T old = *p;
    *p = old + T(v);
    return old;
}
template <class T, class U>
inline T atomicMax(T* p, U v) {
    const T old = *p;
    if (T(v) > old) *p = T(v);
    return old;
}
template <class T, class U>
inline T atomicOr(T* p, U v) {
    const T old = *p;
    *p = old | T(v);
    return old;
}
using std::max;
using std::min;

// Runs f() as one workgroup: here a single thread.
template <class F>
void emu_launch(unsigned, F f) {
    blockDim.x = 1;
    threadIdx.x = 0;
    f();
}
