// Structures shared by the HIP kernels and the host-side runtime of the C ABI.
#pragma once

#include <cstdint>

namespace awcm {

// One independent 1-D solve.  Device-resident; the kernel updates the
// time-loop fields (t, cnt, steps, ...) in place.
struct Inst1D {
    int L, nxc, Nx;            // MRA: levels, coarse points, finest points
    int n_print;
    double gamma, dx, tol, cfl, t_final;
    const double* print_times; // device, n_print entries
    double* Q;                 // 3 planes of Nx+6 doubles (SoA), ghost 3
    double* Qc;                // Q at the start of the step (RK copy)
    double* Qt;                // stage input (stencil source)
    double* F;                 // 3 planes of Nx+1 face fluxes
    uint8_t* masks;            // 5 x (Nx+6): mask_temp, mask0..mask3
    uint8_t* levels;           // 4 x Nx: level0..level3
    double t;                  // time
    double last_dt;
    long long sum_active0;     // sum over steps of |mask0| (throughput numerator)
    int cnt;                   // print counter (1-based, as the drivers)
    int steps;
    int status;                // 0 ok, <0 error
    int pad;
};

}  // namespace awcm
