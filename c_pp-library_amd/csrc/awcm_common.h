// Device-side helpers shared by the 1-D and 2-D AWCM kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "awcm/physics.h"

namespace awcm {

constexpr int kGhost = 3;   // ghost width of every state and mask field

// Declaration keyword of a kernel's dynamic LDS array.  The host emulations
// of the kernels (tests/native/emu*) define it before including this file.
#ifndef AWCM_DYNAMIC_LDS
#define AWCM_DYNAMIC_LDS extern __shared__
#endif
constexpr int kMaxLevels = 16;

// Max of a non-negative double over a wavefront, exact (max is order-free).
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = t_max(v, w);
    }
    return v;
}

// atomicMax on the bit pattern of a non-negative double: IEEE ordering of
// non-negative doubles equals the unsigned ordering of their bits.
__device__ __forceinline__ void atomic_max_pos(unsigned long long* addr, double v) {
    atomicMax(addr, static_cast<unsigned long long>(__double_as_longlong(v)));
}

}  // namespace awcm
