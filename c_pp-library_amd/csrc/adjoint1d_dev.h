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
// the stage input).  The adjoint applies the transposes in reverse order.  The
// trajectory is recomputed segment by segment from checkpoints (a segment's
// stage inputs X0..X2 are kept in HBM while it is reversed), so memory grows
// with the number of checkpoints, not the number of steps.
//
// All transposed scatters are written as gathers: every output element sums
// its contributions in a fixed order, so the adjoint is deterministic (bit-
// identical between runs and independent of the recompute segment length).

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

// Coefficient of the flux at i + o d in dF at coarse point i (coarse_dF,
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

// Adjoint of  Y = a Qc + b (X - h dF(X))  on `mask` (Y = X elsewhere), then of
// the ghost fill that produced X.  Yb: adjoint of Y (ghost entries zero);
// writes Xb (adjoint of the state before the ghost fill) and adds a Yb into
// Cb.  X is the stage input (ghosts filled).
__device__ void stage_T(const Ctx1& c, const Adj1DArgs& A, const double* X, const uint8_t* mask, const uint8_t* lvl,
                        double alpha, double dt, double a, double b, const double* Yb, double* Xb) {
    const int Nx = c.g.Nx, NP = c.NP, maxlev = c.g.maxlev;
    double* Gb = A.Gb;
    double* Cb = A.Cb;
    // per point: adjoint of dF (G), of the copy, and the pass-through part of X
    for (int ii = threadIdx.x; ii < NP; ii += blockDim.x) {
        const int i = ii - G;
        const bool in = i >= 0 && i < Nx;
        const bool on = in && mask[ii];
        const double h = on ? dt / (c.dx * c.g.dxl(lvl[i])) : 0.0;
        for (int k = 0; k < 3; k++) {
            const double y = in ? Yb[k * NP + ii] : 0.0;
            if (on) {
                Gb[k * NP + ii] = -b * h * y;
                Cb[k * NP + ii] += a * y;
                Xb[k * NP + ii] = b * y;
            } else {
                Gb[k * NP + ii] = 0.0;
                Xb[k * NP + ii] = y;
            }
        }
    }
    __syncthreads();
    // faces f = 0..Nx (between f-1 and f): dF_i = F_{i+1} - F_i at finest points
    for (int f = threadIdx.x; f <= Nx; f += blockDim.x) {
        const bool lw = (f >= 1) && mask[f - 1 + G] && lvl[f - 1] == maxlev;
        const bool rw = (f <= Nx - 1) && mask[f + G] && lvl[f] == maxlev;
        double sb[6][3] = {};
        if (lw || rw) {
            double phi[3], s[6][3];
            for (int k = 0; k < 3; k++)
                phi[k] = (lw ? Gb[k * NP + f - 1 + G] : 0.0) - (rw ? Gb[k * NP + f + G] : 0.0);
            for (int p = 0; p < 6; p++)
                for (int k = 0; k < 3; k++) s[p][k] = X[k * NP + f - 3 + p + G];
            face_flux1_vjp(c.gamma, s, alpha, phi, sb);
        }
        double* o = A.FV + (long long)f * 18;
        for (int p = 0; p < 6; p++)
            for (int k = 0; k < 3; k++) o[p * 3 + k] = sb[p][k];
    }
    __syncthreads();
    // gather onto every stencil point (ghosts included)
    for (int ii = threadIdx.x; ii < NP; ii += blockDim.x) {
        const int j = ii - G;
        double acc[3] = {0, 0, 0};
        for (int p = 0; p < 6; p++) {
            const int f = j + 3 - p;
            if (f < 0 || f > Nx) continue;
            for (int k = 0; k < 3; k++) acc[k] += A.FV[(long long)f * 18 + p * 3 + k];
        }
        double Fb[3] = {0, 0, 0};
        bool any = false;
        for (int lev = 0; lev < maxlev; lev++) {
            const int d = c.g.dxl(lev);
            for (int o = -2; o <= 2; o++) {
                const int i = j - o * d;
                if (i < 0 || i >= Nx || !mask[i + G] || lvl[i] != lev) continue;
                const double w = coarse_coef(i, d, Nx - 1, o);
                if (w == 0.0) continue;
                for (int k = 0; k < 3; k++) Fb[k] += w * Gb[k * NP + i + G];
                any = true;
            }
        }
        if (any) {
            double q[3] = {X[ii], X[NP + ii], X[2 * NP + ii]};
            flux_f1_vjp(c.gamma, q, Fb, acc);
        }
        for (int k = 0; k < 3; k++) Xb[k * NP + ii] += acc[k];
    }
    __syncthreads();
    // ghost fill: ghosts copy points 0 and Nx-1 (zero-order extrapolation)
    for (int t = threadIdx.x; t < 3; t += blockDim.x) {
        double* x = Xb + t * NP;
        x[G] += x[G - 1] + x[G - 2] + x[G - 3];
        x[G - 1] = x[G - 2] = x[G - 3] = 0.0;
        x[Nx - 1 + G] += x[Nx + G] + x[Nx + 1 + G] + x[Nx + 2 + G];
        x[Nx + G] = x[Nx + 1 + G] = x[Nx + 2 + G] = 0.0;
    }
    __syncthreads();
}

// Transpose of PrepareEuler with frozen masks, in place on B (3 planes):
// density through IWT^T (mask3) then FWT^T (mask_temp); momentum and energy
// through the interpolation transpose (points of mask3 not in mask_temp).
__device__ void prepare_T(const Ctx1& c, double* B) {
    const int NP = c.NP;
    double* B0 = B;
    double* B1 = B + NP;
    double* B2 = B + 2 * NP;
    const uint8_t* mt = c.mt;
    const uint8_t* m3 = c.m[3];
    // interpolation set, built as InterpolateEuler builds it
    for (int i = threadIdx.x; i < NP; i += blockDim.x) c.tmp[i] = m3[i];
    __syncthreads();
    for (int i = threadIdx.x; i < c.g.Nx; i += blockDim.x)
        if (mt[i + G]) c.tmp[i + G] = 0;
    __syncthreads();
    check_recon(c, c.tmp);
    const uint8_t* S = c.tmp;
    for (int lev = c.g.maxlev - 1; lev >= 0; lev--) {
        const Lift L = lift(c.g, lev);
        // IWT predict^T and interpolation^T: even targets gather from odd sources
        for (int n = threadIdx.x; n <= L.N; n += blockDim.x) {
            double a0 = 0, a1 = 0, a2 = 0;
            const int mlo = max(0, n - 3), mhi = min(L.N - 1, n + 2);
            for (int m = mlo; m <= mhi; m++) {
                int n0, kc;
                odd_stencil(m, L.N, n0, kc);
                const int cc = n - n0;
                if (cc < 0 || cc > 3) continue;
                const int o = (2 * m + 1) * L.h + G;
                const double w = c.W[cc][kc];
                if (m3[o]) a0 += w * B0[o];
                if (S[o] && !mt[o]) { a1 += w * B1[o]; a2 += w * B2[o]; }
            }
            const int e = n * L.d + G;
            B0[e] += a0;
            B1[e] += a1;
            B2[e] += a2;
        }
        __syncthreads();
        for (int m = threadIdx.x; m < L.N; m += blockDim.x) {
            const int o = (2 * m + 1) * L.h + G;
            B0[o] = m3[o] ? 2.0 * B0[o] : 0.0;
            if (S[o] && !mt[o]) { B1[o] = 0.0; B2[o] = 0.0; }
        }
        __syncthreads();
        // IWT update^T: odd targets gather from even sources (n = 1..N-1)
        for (int m = threadIdx.x; m < L.N; m += blockDim.x) {
            double acc = 0;
            const int nlo = max(1, m - 2), nhi = min(L.N - 1, m + 3);
            for (int n = nlo; n <= nhi; n++) {
                int m0, kc;
                even_stencil(n, L.N, m0, kc);
                const int cc = m - m0;
                if (cc < 0 || cc > 3) continue;
                const int e = n * L.d + G;
                if (m3[e]) acc += c.W[cc][kc] * B0[e];
            }
            B0[(2 * m + 1) * L.h + G] -= acc;
        }
        __syncthreads();
        for (int n = 1 + threadIdx.x; n <= L.N - 1; n += blockDim.x) {
            const int e = n * L.d + G;
            if (!m3[e]) B0[e] = 0.0;
        }
        __syncthreads();
    }
    for (int lev = 0; lev < c.g.maxlev; lev++) {
        const Lift L = lift(c.g, lev);
        // FWT update^T: odd targets gather from even sources
        for (int m = threadIdx.x; m < L.N; m += blockDim.x) {
            double acc = 0;
            const int nlo = max(1, m - 2), nhi = min(L.N - 1, m + 3);
            for (int n = nlo; n <= nhi; n++) {
                int m0, kc;
                even_stencil(n, L.N, m0, kc);
                const int cc = m - m0;
                if (cc < 0 || cc > 3) continue;
                const int e = n * L.d + G;
                if (mt[e]) acc += c.W[cc][kc] * B0[e];
            }
            B0[(2 * m + 1) * L.h + G] += acc;
        }
        __syncthreads();
        // FWT predict^T: even targets gather from odd sources; then the 1/2
        for (int n = threadIdx.x; n <= L.N; n += blockDim.x) {
            double acc = 0;
            const int mlo = max(0, n - 3), mhi = min(L.N - 1, n + 2);
            for (int m = mlo; m <= mhi; m++) {
                int n0, kc;
                odd_stencil(m, L.N, n0, kc);
                const int cc = n - n0;
                if (cc < 0 || cc > 3) continue;
                const int o = (2 * m + 1) * L.h + G;
                if (mt[o]) acc += c.W[cc][kc] * B0[o];
            }
            B0[n * L.d + G] -= 0.5 * acc;
        }
        __syncthreads();
        for (int m = threadIdx.x; m < L.N; m += blockDim.x) {
            const int o = (2 * m + 1) * L.h + G;
            if (mt[o]) B0[o] *= 0.5;
        }
        __syncthreads();
    }
}

// Adjoint of tape step n: A.Qb holds the adjoint of Q_{n+1} on entry and of
// Q_n on return.  X holds the step's three stage inputs.
__device__ void step_T(const Ctx1& c, const Adj1DArgs& A, int n, const double* X, double* sc) {
    tape_load(c, A, n, sc);
    const int NP = c.NP, Nx = c.g.Nx;
    const double dt = sc[0];
    // RemoveResidue^T; ghosts of Q_{n+1} are never read by the next step
    for (int ii = threadIdx.x; ii < NP; ii += blockDim.x) {
        const int i = ii - G;
        const bool keep = i >= 0 && i < Nx && c.m[0][ii];
        for (int k = 0; k < 3; k++) {
            if (!keep) A.Qb[k * NP + ii] = 0.0;
            A.Cb[k * NP + ii] = 0.0;
        }
    }
    __syncthreads();
    stage_T(c, A, X + 6 * NP, c.m[0], c.lv[0], sc[3], dt, 1.0 / 3.0, 2.0 / 3.0, A.Qb, A.Xb);
    stage_T(c, A, X + 3 * NP, c.m[1], c.lv[1], sc[2], dt, 0.75, 0.25, A.Xb, A.Qb);
    stage_T(c, A, X, c.m[2], c.lv[2], sc[1], dt, 0.0, 1.0, A.Qb, A.Xb);
    // the RK copy is the first stage input; its adjoint holds no ghost entries,
    // so adding it after the ghost-fill transpose is the same as before
    for (int i = threadIdx.x; i < 3 * NP; i += blockDim.x) A.Qb[i] = A.Xb[i] + A.Cb[i];
    __syncthreads();
    prepare_T(c, A.Qb);
}

}  // namespace

// mode 0: replay the n_steps taped steps from the instance's Q (frozen forward).
// mode 1: reverse them: A.Qb holds dJ/dQ after the last step on entry and
//         dJ/dQ before the first on return; Q is used as work space.
// One workgroup; dynamic LDS 11 * nxp bytes, nxp >= Nx + 6.
__global__ void __launch_bounds__(1024) k1_adjoint(Inst1D* inst, Adj1DArgs A, int mode, int nxp) {
    AWCM_DYNAMIC_LDS uint8_t sm[];
    __shared__ double W[4][5];
    __shared__ double red[16];
    __shared__ double sc[4];
    __shared__ int s_flag;
    if (threadIdx.x == 0) { wavelet_weights(W); s_flag = 0; }
    Inst1D& in = *inst;
    Ctx1 c;
    c.g.L = in.L; c.g.maxlev = in.L - 1; c.g.Nx = in.Nx; c.g.nxc = in.nxc;
    c.gamma = in.gamma; c.dx = in.dx;
    c.NP = in.Nx + 2 * G;
    c.Q = in.Q; c.Qc = in.Qc; c.Qt = in.Qt; c.F = in.F;
    c.mt = sm;
    for (int k = 0; k < 4; k++) c.m[k] = sm + (k + 1) * nxp;
    c.snap = sm + 5 * nxp;
    c.tmp = sm + 6 * nxp;
    for (int k = 0; k < 4; k++) c.lv[k] = sm + (7 + k) * nxp;
    c.W = W; c.red = red; c.flag = &s_flag;
    c.rec = nullptr;
    __syncthreads();
    if (mode == 0) {
        for (int n = 0; n < A.n_steps; n++) replay_step(c, A, n, nullptr, sc);
    } else {
        const int K = A.seg_steps;
        const long long SX = 9LL * c.NP;  // doubles per step in the segment buffer
        for (int s = (A.n_steps + K - 1) / K - 1; s >= 0; s--) {
            const int n0 = s * K, n1 = min(n0 + K, A.n_steps);
            copy_planes(c, c.Q, A.ckpt + (long long)(n0 / A.ckpt_every) * 3 * c.NP);
            __syncthreads();
            for (int n = n0; n < n1; n++) replay_step(c, A, n, A.seg + (n - n0) * SX, sc);
            for (int n = n1 - 1; n >= n0; n--) step_T(c, A, n, A.seg + (n - n0) * SX, sc);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_flag) in.status = -5;
}

}  // namespace awcm

