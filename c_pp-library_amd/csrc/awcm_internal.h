// Structures shared by the HIP kernels and the host-side runtime of the C ABI.
#pragma once

#include <cstddef>
#include <cstdint>

namespace awcm {

// k1_solve mode flag: the instance's state, RK copies and face fluxes live in
// the workgroup's LDS for the launch, after the masks (single-instance
// launches whose arrays fit), copied in at the start and out at the end.
constexpr int kModeLdsState = 0x100;
// LDS bytes of that state for row pitch nxp: Q, Qc, Qt (3 planes each) and
// the face fluxes (3 planes), nxp doubles a plane
constexpr size_t lds_state_bytes(int nxp) { return sizeof(double) * 12 * size_t(nxp); }
// offset of the state in the dynamic LDS: after the 11 * nxp mask bytes
constexpr size_t lds_state_offset(int nxp) { return (size_t(11) * nxp + 15) / 16 * 16; }
// k1_solve mode flag: Q alone lives in LDS after the masks (single-instance
// launches whose full state does not fit but Q does: the transforms, mask
// passes and stage updates then touch LDS, the stage copies and face fluxes
// HBM), copied in at the start and out at the end.
constexpr int kModeLdsQ = 0x200;
constexpr size_t lds_q_bytes(int nxp) { return sizeof(double) * 3 * size_t(nxp); }
// k1_solve mode flag for launches of one workgroup per CU, whose dynamic LDS
// is the whole budget (mode >> 16, in units of 16 bytes): each instance in
// turn keeps its state (planes of its own Nx + 6 doubles) or Q alone in the
// LDS beside the masks when it fits there.  The mode proper is mode & 0xff.
constexpr int kModeLdsAuto = 0x400;
constexpr int kModeBase = 0xff;

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
    // Trajectory tape for the discrete adjoint (null when not recording).  Step
    // n of the recording stores, at tape_m + n * tape_stride, the masks
    // mask_temp, mask0..mask3 (5 x (Nx+6) bytes), level0..level2 (3 x Nx
    // bytes) and the interpolation set of its PrepareEuler (Nx+6 bytes: the
    // points of CheckReconstruction(mask3 minus mask_temp)) that the step
    // used, at tape_s + 4n its dt and the three stage alphas, and every
    // ckpt_every steps the state Q at the step's start (3 x (Nx+6) doubles) at
    // ckpt + (n / ckpt_every) * 3 * (Nx+6).
    uint8_t* tape_m;
    double* tape_s;
    double* ckpt;
    long long tape_stride;
    int tape_n;                // steps recorded so far
    int tape_cap;              // steps the tape can hold
    int ckpt_every;
    int pad2;
};

// Arguments of the 1-D adjoint kernel: the tape of a recorded trajectory
// (layout as in Inst1D) and the adjoint work buffers, all device memory.
struct Adj1DArgs {
    const uint8_t* tape_m;
    const double* tape_s;
    const double* ckpt;
    long long tape_stride;
    int n_steps;      // recorded steps to replay / reverse
    int ckpt_every;   // spacing of the recorded checkpoints
    int seg_steps;    // steps per recompute segment (a multiple of ckpt_every)
    int pad;
    double* seg;      // seg_steps x 3 stage inputs x 3 planes x (Nx+6)
    const int* step;  // the step being reversed (device scalar)
};

// Byte size of one step of the adjoint tape for Nx finest points (masks,
// levels and interpolation set, rounded up to 256 bytes).
inline long long adj_tape_stride(int Nx) {
    const long long NP = Nx + 6;
    return (6 * NP + 3LL * Nx + 255) / 256 * 256;
}

}  // namespace awcm
