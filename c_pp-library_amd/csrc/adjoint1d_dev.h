// Discrete adjoint of the 1-D AWCM Euler step on gfx950, with the per-step
// adaptive data frozen: masks mask_temp/mask0..3, levels level0..2, dt and the
// three Lax-Friedrichs alphas are taken from a tape written by a recording
// forward run.  With them frozen, one step Q_n -> Q_{n+1} is
//
//   P  = PrepareEuler(Q_n)      linear: density FWT (mask_temp), IWT (mask3);
//                               momentum/energy interpolation at new points
//   X0 = ghost fill(P),  Y1 = X0 - h dF(X0)                 on mask2
//   X1 = ghost fill(Y1), Y2 = 3/4 X0 + 1/4 (X1 - h dF(X1))  on mask1
//   X2 = ghost fill(Y2), Y3 = (X0 + 2 (X2 - h dF(X2))) / 3  on mask0
//   Q_{n+1} = RemoveResidue(Y3)                            (zero off mask0)
//
// (AWCM_1D.h:86-120, main_1d_euler.cpp:61-260; points off a stage's mask keep
// the stage input).  This file holds the forward side of the adjoint, one
// workgroup with the masks in LDS: the frozen replay of the whole tape, and
// the recompute of one segment of steps from its checkpoint, which keeps the
// segment's stage inputs X0..X2 in HBM for the reverse sweep
// (adjoint1d_rev.h).

#pragma once

#include "solver1d_dev.h"
#include "awcm/adjoint.h"

namespace awcm {
namespace {

// Lifting level lev: odd points o_m = (2m+1) h (m = 0..N-1), even points
// e_n = n d (n = 0..N), d = 2h = MRA_GRID dx of level lev (WAVELET_1D.h:282-391).
struct Lift {
    int h, d, N;
};
__device__ Lift lift(const Geo1& g, int lev) {
    Lift l;
    l.d = g.dxl(lev);
    l.h = l.d / 2;
    l.N = (g.Nx - 1) / l.d;
    return l;
}
// Stencil of the odd (predict) step at odd point m: even points n0..n0+3,
// weights W[c][kc] (the k = 0 / 1 / 2 cases of the transforms).
__device__ __forceinline__ void odd_stencil(int m, int N, int& n0, int& kc) {
    if (m == 0) n0 = 0, kc = 1;
    else if (m == N - 1) n0 = N - 3, kc = 3;
    else n0 = m - 1, kc = 2;
}
// Stencil of the even (update) step at even point n (1..N-1): odd points m0..m0+3.
__device__ __forceinline__ void even_stencil(int n, int N, int& m0, int& kc) {
    if (n == 1) m0 = 0, kc = 1;
    else if (n == N - 1) m0 = N - 4, kc = 3;
    else m0 = n - 2, kc = 2;
}

// Coefficient of the flux at i + o d in dF at coarse point i (coarse_dF_at,
// main_1d_euler.cpp:111-133, AWCM_1D.h:142-189); 0 if the branch skips it.
__device__ __forceinline__ double pick5(int o, double m2, double m1, double z, double p1, double p2) {
    return o == -2 ? m2 : o == -1 ? m1 : o == 0 ? z : o == 1 ? p1 : p2;
}
__device__ __forceinline__ double coarse_coef(int i, int d, int ie, int o) {
    if (i - 2 * d > 0 && i + 2 * d < ie) return pick5(o, -kD2, -kD1, 0.0, kD1, kD2);
    if (i - 2 * d < 0) return pick5(o, 0.0, kD01, kD11, kD21, kD31);
    if (i == 2 * d) return pick5(o, kD02, kD12, kD22, kD32, kD2);
    if (i > ie - 2 * d) return pick5(o, kD31, kD21, kD11, kD01, 0.0);
    return pick5(o, kD2, kD32, kD22, kD12, kD02);
}

// Load tape step n into the LDS masks and levels of the context.
__device__ void tape_load(const Ctx1& c, const Adj1DArgs& A, int n, double* sc) {
    const uint8_t* src = A.tape_m + n * A.tape_stride;
    for (int i = threadIdx.x; i < c.NP; i += blockDim.x) {
        c.mt[i] = src[i];
        for (int k = 0; k < 4; k++) c.m[k][i] = src[(k + 1) * c.NP + i];
    }
    for (int i = threadIdx.x; i < c.g.Nx; i += blockDim.x)
        for (int k = 0; k < 3; k++) c.lv[k][i] = src[5 * c.NP + k * c.g.Nx + i];
    for (int t = threadIdx.x; t < 4; t += blockDim.x) sc[t] = A.tape_s[4 * n + t];
    __syncthreads();
}

__device__ void copy_planes(const Ctx1& c, double* dst, const double* src) {
    for (int i = threadIdx.x; i < 3 * c.NP; i += blockDim.x) dst[i] = src[i];
}

// Step n of the tape with everything frozen; the arithmetic is the forward
// step's (the FWT consults mask_temp at every point, exactly as the forward
// does before its thresholding touches that point).  If X is given, the three
// stage inputs (after the ghost fill) are stored there.
__device__ void replay_step(const Ctx1& c, const Adj1DArgs& A, int n, double* X, double* sc) {
    tape_load(c, A, n, sc);
    const double dt = sc[0];
    fwt(c, c.mt, -1.0);
    iwt(c, c.m[3]);
    interp(c, c.mt, c.m[3]);
    bc(c);
    if (X) { copy_planes(c, X, c.Q); __syncthreads(); }
    rk_stage(c, 0, c.m[2], c.lv[2], sc[1], dt);
    bc(c);
    if (X) { copy_planes(c, X + 3 * c.NP, c.Q); __syncthreads(); }
    rk_stage(c, 1, c.m[1], c.lv[1], sc[2], dt);
    bc(c);
    if (X) { copy_planes(c, X + 6 * c.NP, c.Q); __syncthreads(); }
    rk_stage(c, 2, c.m[0], c.lv[0], sc[3], dt);
    post(c);
}

}  // namespace

// mode 0: replay the n_steps taped steps from the instance's Q (frozen forward).
// mode 2: if the step *A.step is the last of its recompute segment (segments
//         of A.seg_steps steps from step 0; the tape's last step ends the
//         last one), replay the segment from its checkpoint and keep each
//         step's three stage inputs in A.seg (slot = step - segment start);
//         otherwise return at once.  Q is used as work space.
// One workgroup; dynamic LDS 11 * nxp bytes, nxp >= Nx + 6, and with `lq` the
// work state Q after them (lds_state_offset(nxp) + lds_q_bytes(nxp) bytes),
// copied out to the instance's Q at the end.
__global__ void __launch_bounds__(1024) k1_adjoint(Inst1D* inst, Adj1DArgs A, int mode, int nxp, int lq) {
    AWCM_DYNAMIC_LDS uint8_t sm[];
    __shared__ double W[4][5];
    __shared__ double red[16];
    __shared__ double sc[4];
    __shared__ int s_flag;
    int n0 = 0, n1 = A.n_steps;
    if (mode == 2) {
        const int n = *A.step, K = A.seg_steps;
        if (!(n >= 0 && n < A.n_steps) || !(n == A.n_steps - 1 || n % K == K - 1)) return;
        n0 = n - n % K;
        n1 = n + 1;
    }
    if (threadIdx.x == 0) { wavelet_weights(W); s_flag = 0; }
    Inst1D& in = *inst;
    Ctx1 c;
    c.g.L = in.L; c.g.maxlev = in.L - 1; c.g.Nx = in.Nx; c.g.nxc = in.nxc;
    c.gamma = in.gamma; c.dx = in.dx;
    c.NP = in.Nx + 2 * G;
    c.Q = lq ? reinterpret_cast<double*>(sm + lds_state_offset(nxp)) : in.Q;
    c.Qc = in.Qc; c.Qt = in.Qt; c.F = in.F;
    c.mt = sm;
    for (int k = 0; k < 4; k++) c.m[k] = sm + (k + 1) * nxp;
    c.snap = sm + 5 * nxp;
    c.tmp = sm + 6 * nxp;
    for (int k = 0; k < 4; k++) c.lv[k] = sm + (7 + k) * nxp;
    c.W = W; c.red = red; c.flag = &s_flag;
    c.rec = nullptr;
    __syncthreads();
    if (mode == 0) {
        if (lq) {
            copy_planes(c, c.Q, in.Q);
            __syncthreads();
        }
        for (int n = 0; n < A.n_steps; n++) replay_step(c, A, n, nullptr, sc);
    } else {
        const long long SX = 9LL * c.NP;  // doubles per step in the segment buffer
        copy_planes(c, c.Q, A.ckpt + (long long)(n0 / A.ckpt_every) * 3 * c.NP);
        __syncthreads();
        for (int n = n0; n < n1; n++) replay_step(c, A, n, A.seg + (n - n0) * SX, sc);
    }
    __syncthreads();
    if (lq) copy_planes(c, in.Q, c.Q);
    if (threadIdx.x == 0 && s_flag) in.status = -5;
}

}  // namespace awcm
