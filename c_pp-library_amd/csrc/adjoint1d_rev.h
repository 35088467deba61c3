// Reverse sweep of the discrete adjoint of the 1-D AWCM Euler step (the step
// is written out in adjoint1d_dev.h).  One recorded step n is reversed by a
// chain of grid-wide kernels, each a gather over independent outputs (one
// point, face or lattice node per thread, no LDS, no barriers):
//
//   per stage s = 2, 1, 0 (the RK stages in reverse, on masks 0, 1, 2):
//     k1r_point   adjoint of the stage update at every point: the part that
//                 passes through, the RK-copy part (into Cb) and the adjoint
//                 of dF (Gb); stage 2 first applies RemoveResidue^T
//     k1r_face    vector-Jacobian product of every WENO-Z/Roe face flux a
//                 finest-level point of the stage uses (18 numbers per face)
//     k1r_gather  every point sums the face products and the coarse-level
//                 central-difference terms of its stencil neighbours, then the
//                 ghost-fill transpose folds the ghosts into the end points
//   PrepareEuler^T, per level: the inverse transform's two passes (with the
//   interpolation of momentum and energy) finest to coarsest, then the
//   forward transform's three passes coarsest to finest.
//
// Every output sums its terms in a fixed order, so the sweep is
// deterministic (bit-identical between runs, and independent of the
// recompute segment length).  The step being reversed is read from a device
// scalar, so one step's launch sequence can be captured once as a HIP graph
// and replayed for every step.

#pragma once

#include "adjoint1d_dev.h"

namespace awcm {

// Arguments of the reverse kernels (by value).
struct AdjRev {
    int Nx, L, NP, K;        // finest points, levels, plane pitch, steps per segment
    double gamma, dx;
    const uint8_t* tape_m;   // tape (masks, levels, interpolation set per step)
    long long stride;
    const double* tape_s;    // dt and the three alphas per step
    const double* seg;       // stage inputs of the current segment, K x 9 x NP
    double *Qb, *Xb, *Cb, *Gb, *FV;
    const int* step;         // the step being reversed
};

namespace {

// The 4-point interpolating-wavelet weights W[c][kc], kc = 1, 2, 3 (the rows
// wavelet_weights builds for j = 0, 1, 2; exactly dyadic in FP64).
__device__ __forceinline__ double wlift(int cc, int kc) {
    const double t[3][4] = {{5.0 / 16, 15.0 / 16, -5.0 / 16, 1.0 / 16},
                            {-1.0 / 16, 9.0 / 16, 9.0 / 16, -1.0 / 16},
                            {1.0 / 16, -5.0 / 16, 15.0 / 16, 5.0 / 16}};
    return t[kc - 1][cc];
}

__device__ __forceinline__ int rev_idx() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Per-step views of the tape.  Mask k: 0 mask_temp, 1..4 mask0..mask3.
struct StepView {
    int n, slot;
    const uint8_t* tm;
    const double* sc;
    __device__ const uint8_t* mask(const AdjRev& a, int k) const { return tm + (long long)k * a.NP; }
    __device__ const uint8_t* level(const AdjRev& a, int k) const { return tm + 5LL * a.NP + (long long)k * a.Nx; }
    __device__ const uint8_t* interp_set(const AdjRev& a) const { return tm + 5LL * a.NP + 3LL * a.Nx; }
    __device__ const double* X(const AdjRev& a, int s) const { return a.seg + ((long long)slot * 9 + 3LL * s) * a.NP; }
};
__device__ __forceinline__ StepView step_view(const AdjRev& a) {
    StepView v;
    v.n = *a.step;
    v.slot = v.n % a.K;
    v.tm = a.tape_m + (long long)v.n * a.stride;
    v.sc = a.tape_s + 4LL * v.n;
    return v;
}

// Stage S (2, 1, 0) reads mask0/1/2 with level0/1/2, alpha tape_s[3/2/1],
// combines Y = a Qc + b (X - h dF) with (a, b) = (1/3, 2/3), (3/4, 1/4),
// (0, 1), and maps the adjoint of its output Yb to the adjoint Xb of its
// input: Qb -> Xb (S = 2), Xb -> Qb (S = 1), Qb -> Xb (S = 0).
template <int S> struct StageOf {
    static constexpr int mask = 3 - S;   // mask0 is mask index 1
    static constexpr int level = 2 - S;
    static constexpr int alpha = S + 1;  // tape_s slot of the stage's alpha
};
template <int S> __device__ __forceinline__ double stage_a() { return S == 2 ? 1.0 / 3.0 : S == 1 ? 0.75 : 0.0; }
template <int S> __device__ __forceinline__ double stage_b() { return S == 2 ? 2.0 / 3.0 : S == 1 ? 0.25 : 1.0; }
template <int S> __device__ __forceinline__ const double* stage_in(const AdjRev& a) { return S == 1 ? a.Xb : a.Qb; }
template <int S> __device__ __forceinline__ double* stage_out(const AdjRev& a) { return S == 1 ? a.Qb : a.Xb; }

// Adjoint of the stage update at every ghosted point ii (0 .. NP-1).
template <int S>
__global__ void __launch_bounds__(256) k1r_point(AdjRev a) {
    const int ii = rev_idx();
    if (ii >= a.NP) return;
    const StepView v = step_view(a);
    const uint8_t* mask = v.mask(a, StageOf<S>::mask);
    const uint8_t* lvl = v.level(a, StageOf<S>::level);
    const double dt = v.sc[0];
    const double ca = stage_a<S>(), cb = stage_b<S>();
    const double* Yb = stage_in<S>(a);
    double* Xb = stage_out<S>(a);
    const int i = ii - G;
    const bool in = i >= 0 && i < a.Nx;
    // stage 2's mask is mask0, so on/off also applies RemoveResidue^T there
    const bool on = in && mask[ii];
    const double h = on ? dt / (a.dx * double(1 << (a.L - 1 - lvl[i]))) : 0.0;
    for (int k = 0; k < 3; k++) {
        const long long o = (long long)k * a.NP + ii;
        const double y = (S == 2 ? on : in) ? Yb[o] : 0.0;
        if (on) {
            a.Gb[o] = -cb * h * y;
            if (S == 2) a.Cb[o] = 0.0 + ca * y;
            if (S == 1) a.Cb[o] += ca * y;
            Xb[o] = cb * y;
        } else {
            a.Gb[o] = 0.0;
            if (S == 2) a.Cb[o] = 0.0;
            Xb[o] = y;
        }
    }
}

// Product of every face flux f = 0 .. Nx (between points f-1 and f) with the
// adjoint of the two point derivatives it enters, onto its 6 stencil states.
template <int S>
__global__ void __launch_bounds__(256) k1r_face(AdjRev a) {
    const int f = rev_idx();
    if (f > a.Nx) return;
    const StepView v = step_view(a);
    const uint8_t* mask = v.mask(a, StageOf<S>::mask);
    const uint8_t* lvl = v.level(a, StageOf<S>::level);
    const int maxlev = a.L - 1;
    const bool lw = (f >= 1) && mask[f - 1 + G] && lvl[f - 1] == maxlev;
    const bool rw = (f <= a.Nx - 1) && mask[f + G] && lvl[f] == maxlev;
    double sb[6][3] = {};
    if (lw || rw) {
        const double* X = v.X(a, S);
        double phi[3], s[6][3];
        for (int k = 0; k < 3; k++) {
            const double gl = lw ? a.Gb[(long long)k * a.NP + f - 1 + G] : 0.0;
            const double gr = rw ? a.Gb[(long long)k * a.NP + f + G] : 0.0;
            phi[k] = gl - gr;
        }
        for (int p = 0; p < 6; p++)
            for (int k = 0; k < 3; k++) s[p][k] = X[(long long)k * a.NP + f - 3 + p + G];
        face_flux1_vjp(a.gamma, s, v.sc[StageOf<S>::alpha], phi, sb);
    }
    double* o = a.FV + (long long)f * 18;
    for (int p = 0; p < 6; p++)
        for (int k = 0; k < 3; k++) o[p * 3 + k] = sb[p][k];
}

// Sum of every term of the stage's flux derivatives that reads the state at
// ghosted point j (-3 .. Nx+2): the face products of the faces whose stencil
// holds j, and the physical-flux product of the coarse-level differences
// whose stencil holds j.
template <int S>
__device__ void gather_at(const AdjRev& a, const StepView& v, int j, double* acc) {
    const uint8_t* mask = v.mask(a, StageOf<S>::mask);
    const uint8_t* lvl = v.level(a, StageOf<S>::level);
    const int Nx = a.Nx, maxlev = a.L - 1;
    acc[0] = acc[1] = acc[2] = 0.0;
    for (int p = 0; p < 6; p++) {
        const int f = j + 3 - p;
        if (f < 0 || f > Nx) continue;
        for (int k = 0; k < 3; k++) acc[k] += a.FV[(long long)f * 18 + p * 3 + k];
    }
    double Fb[3] = {0, 0, 0};
    bool any = false;
    for (int lev = 0; lev < maxlev; lev++) {
        const int d = 1 << (a.L - 1 - lev);
        for (int o = -2; o <= 2; o++) {
            const int i = j - o * d;
            if (i < 0 || i >= Nx || !mask[i + G] || lvl[i] != lev) continue;
            const double w = coarse_coef(i, d, Nx - 1, o);
            if (w == 0.0) continue;
            for (int k = 0; k < 3; k++) Fb[k] += w * a.Gb[(long long)k * a.NP + i + G];
            any = true;
        }
    }
    if (any) {
        const double* X = v.X(a, S);
        const double q[3] = {X[j + G], X[a.NP + j + G], X[2LL * a.NP + j + G]};
        flux_f1_vjp(a.gamma, q, Fb, acc);
    }
}

// The adjoint of the stage input at every ghosted point: pass-through part
// plus the gathered terms, with the ghost-fill transpose (ghosts copy points 0
// and Nx-1) folded into the end points; ghost entries end at 0.  Stage 0 adds
// the RK copy's adjoint (the copy is the first stage input) and writes Qb.
template <int S>
__global__ void __launch_bounds__(256) k1r_gather(AdjRev a) {
    const int ii = rev_idx();
    if (ii >= a.NP) return;
    double* out = S == 0 ? a.Qb : stage_out<S>(a);
    const double* Xp = stage_out<S>(a);  // pass-through part from k1r_point
    const int j = ii - G;
    if (j < 0 || j >= a.Nx) {
        for (int k = 0; k < 3; k++) out[(long long)k * a.NP + ii] = 0.0;
        return;
    }
    const StepView v = step_view(a);
    double acc[3], gh[3], t[3];
    gather_at<S>(a, v, j, acc);
    double val[3];
    for (int k = 0; k < 3; k++) val[k] = Xp[(long long)k * a.NP + ii] + acc[k];
    if (j == 0 || j == a.Nx - 1) {
        const int g0 = j == 0 ? -1 : a.Nx;
        const int dg = j == 0 ? -1 : 1;
        gather_at<S>(a, v, g0, gh);
        gather_at<S>(a, v, g0 + dg, t);
        for (int k = 0; k < 3; k++) gh[k] += t[k];
        gather_at<S>(a, v, g0 + 2 * dg, t);
        for (int k = 0; k < 3; k++) val[k] += gh[k] + t[k];
    }
    if (S == 0)
        for (int k = 0; k < 3; k++) val[k] += a.Cb[(long long)k * a.NP + ii];
    for (int k = 0; k < 3; k++) out[(long long)k * a.NP + ii] = val[k];
}

// PrepareEuler^T on Qb (planes B0 density, B1 momentum, B2 energy).
// Level lev: odd points o_m = (2m+1) h (m = 0..N-1), even points e_n = n d
// (n = 0..N), d = 2h.

// Inverse transform, even targets: predict^T of the density (mask3) and the
// interpolation^T of momentum and energy (interpolation set minus mask_temp)
// gather from the odd points; then the update's zeroing of even density
// points outside mask3 (n = 1..N-1; the odd pass below reads only even points
// inside mask3, so it may come first).
__device__ __forceinline__ void k1r_iwt_even_at(const AdjRev& a, int lev, int idx) {
    const int d = 1 << (a.L - 1 - lev), hh = d / 2, N = (a.Nx - 1) / d;
    const int n = idx;
    if (n > N) return;
    const StepView v = step_view(a);
    const uint8_t* mt = v.mask(a, 0);
    const uint8_t* m3 = v.mask(a, 4);
    const uint8_t* S = v.interp_set(a);
    double* B0 = a.Qb;
    double* B1 = a.Qb + a.NP;
    double* B2 = a.Qb + 2LL * a.NP;
    double a0 = 0, a1 = 0, a2 = 0;
    const int mlo = max(0, n - 3), mhi = min(N - 1, n + 2);
    for (int m = mlo; m <= mhi; m++) {
        int n0, kc;
        odd_stencil(m, N, n0, kc);
        const int cc = n - n0;
        if (cc < 0 || cc > 3) continue;
        const int o = (2 * m + 1) * hh + G;
        const double w = wlift(cc, kc);
        if (m3[o]) a0 += w * B0[o];
        if (S[o] && !mt[o]) { a1 += w * B1[o]; a2 += w * B2[o]; }
    }
    const int e = n * d + G;
    B0[e] = (n >= 1 && n <= N - 1 && !m3[e]) ? 0.0 : B0[e] + a0;
    B1[e] += a1;
    B2[e] += a2;
}
__global__ void __launch_bounds__(256) k1r_iwt_even(AdjRev a, int lev) { k1r_iwt_even_at(a, lev, rev_idx()); }

// Inverse transform, odd targets: the predict's doubling (points outside
// mask3 end at 0), minus the update^T gathered from the even points inside
// mask3; interpolated momentum and energy points end at 0.
__device__ __forceinline__ void k1r_iwt_odd_at(const AdjRev& a, int lev, int idx) {
    const int d = 1 << (a.L - 1 - lev), hh = d / 2, N = (a.Nx - 1) / d;
    const int m = idx;
    if (m >= N) return;
    const StepView v = step_view(a);
    const uint8_t* mt = v.mask(a, 0);
    const uint8_t* m3 = v.mask(a, 4);
    const uint8_t* S = v.interp_set(a);
    double* B0 = a.Qb;
    const int o = (2 * m + 1) * hh + G;
    double acc = 0;
    const int nlo = max(1, m - 2), nhi = min(N - 1, m + 3);
    for (int n = nlo; n <= nhi; n++) {
        int m0, kc;
        even_stencil(n, N, m0, kc);
        const int cc = m - m0;
        if (cc < 0 || cc > 3) continue;
        const int e = n * d + G;
        if (m3[e]) acc += wlift(cc, kc) * B0[e];
    }
    B0[o] = (m3[o] ? 2.0 * B0[o] : 0.0) - acc;
    if (S[o] && !mt[o]) {
        a.Qb[a.NP + o] = 0.0;
        a.Qb[2LL * a.NP + o] = 0.0;
    }
}
__global__ void __launch_bounds__(256) k1r_iwt_odd(AdjRev a, int lev) { k1r_iwt_odd_at(a, lev, rev_idx()); }

// Forward transform, odd targets: update^T gathered from the even points
// inside mask_temp.
__device__ __forceinline__ void k1r_fwt_odd_at(const AdjRev& a, int lev, int idx) {
    const int d = 1 << (a.L - 1 - lev), hh = d / 2, N = (a.Nx - 1) / d;
    const int m = idx;
    if (m >= N) return;
    const StepView v = step_view(a);
    const uint8_t* mt = v.mask(a, 0);
    double* B0 = a.Qb;
    double acc = 0;
    const int nlo = max(1, m - 2), nhi = min(N - 1, m + 3);
    for (int n = nlo; n <= nhi; n++) {
        int m0, kc;
        even_stencil(n, N, m0, kc);
        const int cc = m - m0;
        if (cc < 0 || cc > 3) continue;
        const int e = n * d + G;
        if (mt[e]) acc += wlift(cc, kc) * B0[e];
    }
    B0[(2 * m + 1) * hh + G] += acc;
}
__global__ void __launch_bounds__(256) k1r_fwt_odd(AdjRev a, int lev) { k1r_fwt_odd_at(a, lev, rev_idx()); }

// Forward transform, even targets: predict^T gathered from the odd points
// inside mask_temp, with the predict's factor 1/2.
__device__ __forceinline__ void k1r_fwt_even_at(const AdjRev& a, int lev, int idx) {
    const int d = 1 << (a.L - 1 - lev), hh = d / 2, N = (a.Nx - 1) / d;
    const int n = idx;
    if (n > N) return;
    const StepView v = step_view(a);
    const uint8_t* mt = v.mask(a, 0);
    double* B0 = a.Qb;
    double acc = 0;
    const int mlo = max(0, n - 3), mhi = min(N - 1, n + 2);
    for (int m = mlo; m <= mhi; m++) {
        int n0, kc;
        odd_stencil(m, N, n0, kc);
        const int cc = n - n0;
        if (cc < 0 || cc > 3) continue;
        const int o = (2 * m + 1) * hh + G;
        if (mt[o]) acc += wlift(cc, kc) * B0[o];
    }
    B0[n * d + G] -= 0.5 * acc;
}
__global__ void __launch_bounds__(256) k1r_fwt_even(AdjRev a, int lev) { k1r_fwt_even_at(a, lev, rev_idx()); }

// Forward transform, odd targets inside mask_temp: the predict's factor 1/2.
__device__ __forceinline__ void k1r_fwt_half_at(const AdjRev& a, int lev, int idx) {
    const int d = 1 << (a.L - 1 - lev), hh = d / 2, N = (a.Nx - 1) / d;
    const int m = idx;
    if (m >= N) return;
    const StepView v = step_view(a);
    const int o = (2 * m + 1) * hh + G;
    if (v.mask(a, 0)[o]) a.Qb[o] *= 0.5;
}
__global__ void __launch_bounds__(256) k1r_fwt_half(AdjRev a, int lev) { k1r_fwt_half_at(a, lev, rev_idx()); }

// Step counter: the next replay reverses the step before.
__global__ void k1r_next(int* step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step -= 1;
}

// The transposed transforms' passes of the coarse levels top..0 (at most
// kRevWg points each) back to back in one workgroup: the inverse
// transform's, coarsest last, then the forward transform's, coarsest first --
// the same gathers as the grid kernels, a barrier after each pass.  One
// launch where there were five a level, each a launch of a few workgroups.
constexpr int kRevWg = 1024;
__global__ void __launch_bounds__(kRevWg) k1r_coarse_levels(AdjRev a, int top) {
    for (int lev = top; lev >= 0; lev--) {
        const int N = (a.Nx - 1) >> (a.L - 1 - lev);
        for (int n = threadIdx.x; n <= N; n += blockDim.x) k1r_iwt_even_at(a, lev, n);
        __syncthreads();
        for (int m = threadIdx.x; m < N; m += blockDim.x) k1r_iwt_odd_at(a, lev, m);
        __syncthreads();
    }
    for (int lev = 0; lev <= top; lev++) {
        const int N = (a.Nx - 1) >> (a.L - 1 - lev);
        for (int m = threadIdx.x; m < N; m += blockDim.x) k1r_fwt_odd_at(a, lev, m);
        __syncthreads();
        for (int n = threadIdx.x; n <= N; n += blockDim.x) k1r_fwt_even_at(a, lev, n);
        __syncthreads();
        for (int m = threadIdx.x; m < N; m += blockDim.x) k1r_fwt_half_at(a, lev, m);
        __syncthreads();
    }
}

}  // namespace

// Launches of one reversed step (the recompute launch comes before them).
// AWCM_REV_LAUNCH(kernel, work items, args...) is a grid launch of 256-thread
// workgroups covering the items; the host emulation of the tests defines it
// as a loop.
#ifndef AWCM_REV_LAUNCH
#define AWCM_REV_LAUNCH(kern, items, ...) \
    hipLaunchKernelGGL((kern), dim3(((items) + 255) / 256), dim3(256), 0, stream, __VA_ARGS__)
#endif
// AWCM_REV_LAUNCH_WG(kernel, args...): one workgroup of kRevWg threads
#ifndef AWCM_REV_LAUNCH_WG
#define AWCM_REV_LAUNCH_WG(kern, ...) hipLaunchKernelGGL((kern), dim3(1), dim3(kRevWg), 0, stream, __VA_ARGS__)
#endif

inline void launch_reverse_step(const AdjRev& a, hipStream_t stream) {
    (void)stream;
    const int maxlev = a.L - 1;
    AWCM_REV_LAUNCH(k1r_point<2>, a.NP, a);
    AWCM_REV_LAUNCH(k1r_face<2>, a.Nx + 1, a);
    AWCM_REV_LAUNCH(k1r_gather<2>, a.NP, a);
    AWCM_REV_LAUNCH(k1r_point<1>, a.NP, a);
    AWCM_REV_LAUNCH(k1r_face<1>, a.Nx + 1, a);
    AWCM_REV_LAUNCH(k1r_gather<1>, a.NP, a);
    AWCM_REV_LAUNCH(k1r_point<0>, a.NP, a);
    AWCM_REV_LAUNCH(k1r_face<0>, a.Nx + 1, a);
    AWCM_REV_LAUNCH(k1r_gather<0>, a.NP, a);
    // levels whose passes fit one workgroup: 0..top
    int top = -1;
    while (top + 1 < maxlev && ((a.Nx - 1) >> (a.L - 2 - top)) + 1 <= kRevWg) top++;
    for (int lev = maxlev - 1; lev > top; lev--) {
        const int N = (a.Nx - 1) >> (a.L - 1 - lev);
        AWCM_REV_LAUNCH(k1r_iwt_even, N + 1, a, lev);
        AWCM_REV_LAUNCH(k1r_iwt_odd, N, a, lev);
    }
    if (top >= 0) AWCM_REV_LAUNCH_WG(k1r_coarse_levels, a, top);
    for (int lev = top + 1; lev < maxlev; lev++) {
        const int N = (a.Nx - 1) >> (a.L - 1 - lev);
        AWCM_REV_LAUNCH(k1r_fwt_odd, N, a, lev);
        AWCM_REV_LAUNCH(k1r_fwt_even, N + 1, a, lev);
        AWCM_REV_LAUNCH(k1r_fwt_half, N, a, lev);
    }
}

}  // namespace awcm
