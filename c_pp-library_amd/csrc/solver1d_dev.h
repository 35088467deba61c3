// Device side of the 1-D AWCM solver: geometry, the mask/transform passes of
// PrepareEuler, the SSP-RK3 stage with WENO-Z/Roe faces, and one driver-loop
// step.  Shared by the forward solve kernel and the adjoint kernel; every
// function runs on one workgroup that owns one instance, with masks and levels
// in LDS and fields in HBM (planes of Nx+6 doubles, SoA).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "awcm_common.h"
#include "awcm_internal.h"

namespace awcm {
namespace {

constexpr int G = kGhost;

// MRA_GRID_INT_1D geometry (MRA_GRID_UNIFORM_1D.h:70-98) for one instance.
// The reference's integer ranges divide by the point counts of a level; with
// Nx - 1 = (nxc - 1) 2^(L-1) those quotients are exact powers of two, so the
// spacing of grid[lev] and grid_x[lev] is 2^(L-1-lev) and a range of n
// points spans (n - 1) << (L-1-lev): shifts, not integer divisions.
struct Geo1 {
    int L, maxlev, Nx, nxc;
    __device__ int shl(int lev) const { return L - 1 - lev; }
    __device__ int dxl(int lev) const { return 1 << (L - 1 - lev); }
    // grid_x[lev]: odd points of level lev+1
    __device__ void gridx(int lev, int& s, int& e, int& d) const {
        s = dxl(lev + 1);
        e = (Nx - 1) - dxl(lev + 1);
        d = dxl(lev);
    }
    // grid[lev]
    __device__ void grid(int lev, int& s, int& e, int& d) const {
        s = 0;
        e = Nx - 1;
        d = dxl(lev);
    }
};

struct Ctx1 {
    Geo1 g;
    double gamma, dx;
    double* Q;    // 3 planes, stride NP = Nx + 6, index c*NP + i + 3
    double* Qc;
    double* Qt;
    double* F;    // 3 planes of Nx + 1 face fluxes
    int NP;
    uint8_t *mt, *m[4], *snap, *tmp;  // LDS, index i + 3
    uint8_t* lv[4];                   // LDS, index i
    double (*W)[5];
    double* red;
    int* flag;    // LDS error flag (out-of-range scatter)
    Inst1D* rec;  // instance whose tape records the steps (null: no recording)
};

__device__ __forceinline__ void set1(const Ctx1& c, uint8_t* m, int i) {
    if (i >= -G && i <= c.g.Nx - 1 + G) m[i + G] = 1;
    else *c.flag = 1;
}

// Block-wide max of a non-negative double (exact).
__device__ double block_max(double v, double* red) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = 0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int k = 0; k < nw; k++) r = t_max(r, red[k]);
    return r;
}

__device__ void copy_mask(const Ctx1& c, uint8_t* dst, const uint8_t* src) {
    for (int i = threadIdx.x; i < c.g.Nx + 2 * G; i += blockDim.x) dst[i] = src[i];
}

// WAVELET_1D::ForwardWaveletTransformEuler (WAVELET_1D.h:282-335): density only.
__device__ void fwt(const Ctx1& c, uint8_t* mask, double tol) {
    double* q = c.Q;
    const int ie = c.g.Nx - 1;
    for (int lev = c.g.maxlev - 1; lev >= 0; lev--) {
        int s, e, d;
        c.g.gridx(lev, s, e, d);
        const int dh = d >> 1;
        const int n = ((e - s) >> c.g.shl(lev)) + 1;
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            const int i = s + t * d;
            if (mask[i + G] == 0) continue;
            int is = i - d - dh, k = 1;
            if (is < 0) is = 0, k = 0;
            else if (i + d + dh > ie) is = ie - 3 * d, k = 2;
            double v = q[i + G];
            for (int cc = 0; cc < 4; cc++) v -= c.W[cc][k + 1] * q[is + cc * d + G];
            v *= 0.5;
            q[i + G] = v;
            if (t_abs(v) < tol) mask[i + G] = 0;
        }
        __syncthreads();
        int gs, ge, gd;
        c.g.grid(lev, gs, ge, gd);
        const int nu = ((ge - gd - (gs + gd)) >> c.g.shl(lev)) + 1;
        for (int t = threadIdx.x; t < nu; t += blockDim.x) {
            const int i = gs + gd + t * gd;
            if (mask[i + G] == 0) continue;
            int is = i - d - dh, k = 1;
            if (is < 0) is = dh, k = 0;
            else if (i + d + dh > ie) is = ie - 3 * d - dh, k = 2;
            double v = q[i + G];
            for (int cc = 0; cc < 4; cc++) v += c.W[cc][k + 1] * q[is + cc * d + G];
            q[i + G] = v;
        }
        __syncthreads();
    }
}

// WAVELET_1D::InverseWaveletTransformEuler (WAVELET_1D.h:337-391)
__device__ void iwt(const Ctx1& c, const uint8_t* mask) {
    double* q = c.Q;
    const int ie = c.g.Nx - 1;
    for (int lev = 0; lev < c.g.maxlev; lev++) {
        int s, e, d;
        c.g.gridx(lev, s, e, d);
        const int dh = d >> 1;
        int gs, ge, gd;
        c.g.grid(lev, gs, ge, gd);
        const int nu = ((ge - gd - (gs + gd)) >> c.g.shl(lev)) + 1;
        for (int t = threadIdx.x; t < nu; t += blockDim.x) {
            const int i = gs + gd + t * gd;
            if (mask[i + G] == 0) { q[i + G] = 0; continue; }
            int is = i - d - dh, k = 1;
            if (is < 0) is = dh, k = 0;
            else if (i + d + dh > ie) is = ie - 3 * d - dh, k = 2;
            double v = q[i + G];
            for (int cc = 0; cc < 4; cc++) v -= c.W[cc][k + 1] * q[is + cc * d + G];
            q[i + G] = v;
        }
        __syncthreads();
        const int n = ((e - s) >> c.g.shl(lev)) + 1;
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            const int i = s + t * d;
            if (mask[i + G] == 0) { q[i + G] = 0; continue; }
            int is = i - d - dh, k = 1;
            if (is < 0) is = 0, k = 0;
            else if (i + d + dh > ie) is = ie - 3 * d, k = 2;
            double v = q[i + G];
            v *= 2;
            for (int cc = 0; cc < 4; cc++) v += c.W[cc][k + 1] * q[is + cc * d + G];
            q[i + G] = v;
        }
        __syncthreads();
    }
}

// AWCM_MANAGER_1D::CheckReconstruction (AWCM_MANAGER_1D.h:45-71): level by level,
// finest first; within a level the writes land on level-lev points only.
__device__ void check_recon(const Ctx1& c, uint8_t* mask) {
    const int ie = c.g.Nx - 1;
    for (int lev = c.g.maxlev - 1; lev >= 0; lev--) {
        int s, e, d;
        c.g.gridx(lev, s, e, d);
        const int dh = d >> 1;
        const int n = ((e - s) >> c.g.shl(lev)) + 1;
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            const int i = s + t * d;
            if (mask[i + G] == 0) continue;
            int is = i - d - dh;
            if (is < 0) is = 0;
            else if (i + d + dh > ie) is = ie - 3 * d;
            for (int cc = 0; cc < 4; cc++) set1(c, mask, is + cc * d);
        }
        __syncthreads();
    }
}

// AWCM_MANAGER_1D::AddAdjacentZone (AWCM_MANAGER_1D.h:14-43), from a snapshot.
__device__ void adjacent(const Ctx1& c, uint8_t* mask) {
    copy_mask(c, c.snap, mask);
    __syncthreads();
    {
        int s, e, d;
        c.g.gridx(c.g.maxlev - 1, s, e, d);
        const int n = ((e - s) >> c.g.shl(c.g.maxlev - 1)) + 1;
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            const int i = s + t * d;
            if (c.snap[i + G]) { set1(c, mask, i - 1); set1(c, mask, i + 1); }
        }
    }
    for (int lev = c.g.maxlev - 2; lev >= 0; lev--) {
        int gs, ge, gd;
        c.g.grid(lev, gs, ge, gd);
        const int dq = gd >> 2;
        int s, e, d;
        c.g.gridx(lev, s, e, d);
        const int n = ((e - s) >> c.g.shl(lev)) + 1;
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            const int i = s + t * d;
            if (c.snap[i + G]) { set1(c, mask, i - dq); set1(c, mask, i + dq); }
        }
    }
    __syncthreads();
}

// AWCM_MANAGER_1D::AddPointForDifferentiation (AWCM_MANAGER_1D.h:73-107)
__device__ void addpoint(const Ctx1& c, uint8_t* mask, const uint8_t* level) {
    copy_mask(c, c.snap, mask);
    __syncthreads();
    const int N = c.g.Nx - 1;
    for (int i = threadIdx.x; i <= N; i += blockDim.x) {
        if (c.snap[i + G] == 0) continue;
        const int lev = level[i];
        int gs, ge, d;
        c.g.grid(lev, gs, ge, d);
        const int is = i - 2 * d, iend = i + 2 * d;
        if (lev == c.g.maxlev) {
            for (int k = -3; k < 4; k++) set1(c, mask, i + k);
        } else if (is <= 0) {
            for (int k = 0; k < 6; k++) set1(c, mask, k * d);
        } else if (iend >= N) {
            for (int k = 0; k < 6; k++) set1(c, mask, N - k * d);
        } else {
            for (int k = 0; k < 5; k++) set1(c, mask, is + k * d);
        }
    }
    __syncthreads();
}

// AWCM_MANAGER_1D::ComputeLevel (AWCM_MANAGER_1D.h:109-133).  A point with no
// active neighbour keeps its previous level, as in the reference.
// The search by spacing d = 1, 2, 4, ... reads every spacing's two mask bytes
// at once (clamped into the ghosted range, so every read is in the array) and
// then takes the first spacing that has an active neighbour, or that leaves
// the ghosted grid (the error flag), as the reference's loop does one
// spacing after the other.  Out of line, like the stage functions, so its
// loads in flight do not add to the solve kernel's register count.
__device__ __attribute__((noinline)) void level_at(int Nx, int maxlev, uint8_t* lvl, const uint8_t* mask, int* flag) {
    const uint32_t spacings = (2u << maxlev) - 1u;
    for (int i = 1 + threadIdx.x; i <= Nx - 2; i += blockDim.x) {
        if (mask[i + G] == 0) { lvl[i] = 0; continue; }
        uint32_t hit = 0, out = 0;
#pragma unroll
        for (int q = 0; q < kMaxLevels; q++) {
            const int d = 1 << q;
            const bool in = i - d >= -G && i + d <= Nx - 1 + G;
            const int lo = in ? i - d : -G, hi = in ? i + d : -G;
            if (!in) out |= 1u << q;
            else if (mask[lo + G] | mask[hi + G]) hit |= 1u << q;
        }
        const uint32_t first = (hit | out) & spacings;
        if (first) {
            const int q = __builtin_ctz(first);
            if ((out >> q) & 1u) *flag = 1;
            else lvl[i] = (uint8_t)(maxlev - q);
        }
    }
    if (threadIdx.x == 0) { lvl[0] = (uint8_t)maxlev; lvl[Nx - 1] = (uint8_t)maxlev; }
    __syncthreads();
}
__device__ void level(const Ctx1& c, uint8_t* lvl, const uint8_t* mask) {
    level_at(c.g.Nx, c.g.maxlev, lvl, mask, c.flag);
}

// WAVELET_1D::InterpolateEuler(q, mra, from, to) (WAVELET_1D.h:393-431): momentum
// and energy at points of `to` that were not active in `from`.
// Components P0..P1-1: 1..2 for InterpolateEuler (WAVELET_1D.h:393-431), 0
// for the scalar Interpolate(q, mra, from, to) (WAVELET_1D.h:198-233).
template <int P0 = 1, int P1 = 3>
__device__ void interp(const Ctx1& c, const uint8_t* from, const uint8_t* to) {
    const int ie = c.g.Nx - 1;
    for (int i = threadIdx.x; i < c.g.Nx + 2 * G; i += blockDim.x) c.tmp[i] = to[i];
    __syncthreads();
    for (int i = threadIdx.x; i <= ie; i += blockDim.x)
        if (from[i + G]) c.tmp[i + G] = 0;
    __syncthreads();
    check_recon(c, c.tmp);
    for (int lev = 0; lev < c.g.maxlev; lev++) {
        int s, e, d;
        c.g.gridx(lev, s, e, d);
        const int dh = d >> 1;
        const int n = ((e - s) >> c.g.shl(lev)) + 1;
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            const int i = s + t * d;
            if (c.tmp[i + G] == 0 || from[i + G] == 1) continue;
            int is = i - d - dh, k = 1;
            if (is < 0) is = 0, k = 0;
            else if (i + d + dh > ie) is = ie - 3 * d, k = 2;
            for (int p = P0; p < P1; p++) {
                double* q = c.Q + p * c.NP;
                double v = 0;
                for (int cc = 0; cc < 4; cc++) v += c.W[cc][k + 1] * q[is + cc * d + G];
                q[i + G] = v;
            }
        }
        __syncthreads();
    }
}

// AWCM_1D::BadingBoundary (AWCM_1D.h:73-80): 3-wide band on both ends.
__device__ void band(const Ctx1& c, uint8_t* mask) {
    const int Nx = c.g.Nx;
    for (int i = threadIdx.x; i < Nx; i += blockDim.x)
        if (!(3 <= i && i <= Nx - 1 - 3)) mask[i + G] = 1;
    __syncthreads();
}

// AWCM_1D::PrepareEuler (AWCM_1D.h:86-120)
__device__ void prepare(const Ctx1& c, double tol) {
    copy_mask(c, c.mt, c.m[0]);
    __syncthreads();
    fwt(c, c.m[0], tol);
    band(c, c.m[0]);
    adjacent(c, c.m[0]);
    check_recon(c, c.m[0]);
    for (int k = 1; k <= 3; k++) {
        copy_mask(c, c.m[k], c.m[k - 1]);
        __syncthreads();
        level(c, c.lv[k - 1], c.m[k - 1]);
        adjacent(c, c.m[k]);
        addpoint(c, c.m[k], c.lv[k - 1]);
        check_recon(c, c.m[k]);
    }
    level(c, c.lv[3], c.m[3]);
    iwt(c, c.m[3]);
    interp(c, c.mt, c.m[3]);
}

// AWCM_1D::Prepare (AWCM_1D.h:45-71) on the scalar field (plane 0): no
// mask_temp, no AddAdjacentZone on masks 1..3, no interpolation.
__device__ void prepare_scalar(const Ctx1& c, double tol) {
    fwt(c, c.m[0], tol);
    band(c, c.m[0]);
    adjacent(c, c.m[0]);
    check_recon(c, c.m[0]);
    for (int k = 1; k <= 3; k++) {
        copy_mask(c, c.m[k], c.m[k - 1]);
        __syncthreads();
        level(c, c.lv[k - 1], c.m[k - 1]);
        addpoint(c, c.m[k], c.lv[k - 1]);
        check_recon(c, c.m[k]);
    }
    level(c, c.lv[3], c.m[3]);
    iwt(c, c.m[3]);
}

// Zero-order extrapolation into the ghosts (PROBLEM_SHOCKTUBE.h:86-107; the
// Shu-Osher manager PROBLEM_ENTROPY_INTERACTING.h:81-102 is the same).
__device__ void bc(const Ctx1& c) {
    const int Nx = c.g.Nx;
    for (int t = threadIdx.x; t < 3 * G; t += blockDim.x) {
        const int p = t / G, k = t % G;
        double* q = c.Q + p * c.NP;
        q[-1 - k + G] = q[0 + G];
        q[Nx + k + G] = q[Nx - 1 + G];
    }
    __syncthreads();
}

// Read point i of a 3-plane field; an index outside the ghosted range (which a
// closed mask never produces) is clamped and reported through the error flag.
__device__ void load3(const Ctx1& c, const double* Q, int i, double* o) {
    if (i < -G || i > c.g.Nx - 1 + G) { *c.flag = 1; i = i < -G ? -G : c.g.Nx - 1 + G; }
    o[0] = Q[i + G];
    o[1] = Q[c.NP + i + G];
    o[2] = Q[2 * c.NP + i + G];
}

// Three doubles returned in registers by the out-of-line stage functions
// below.  Those two hold nearly all of a stage's arithmetic (the WENO face
// fluxes and the coarse-point flux differences); inlined, their registers
// push the rest of the solve kernel past the 128 VGPRs a 1024-thread
// workgroup allows, and every phase then reloads spilled values from scratch
// memory; called, they spill (if at all) only inside themselves.
struct F3 { double v[3]; };

// Flux derivative at a coarse point (main_1d_euler.cpp:111-133; AWCM_1D.h:142-189)
// of the stage input Qt (3 planes of NP doubles); a stencil index outside the
// ghosted range (which a closed mask never produces) is clamped and reported
// through *flag.
__device__ __attribute__((noinline)) F3 coarse_dF_at(const double* Qt, int NP, int Nx, double g, int i, int d,
                                                     int* flag) {
    const int ie = Nx - 1;
    F3 r;
    double* out = r.v;
    auto load3 = [&](const double* Q, int k, double* o) {
        if (k < -G || k > Nx - 1 + G) { *flag = 1; k = k < -G ? -G : Nx - 1 + G; }
        o[0] = Q[k + G];
        o[1] = Q[NP + k + G];
        o[2] = Q[2 * NP + k + G];
    };
    double q[3], fll[3], fl[3], fc[3], fr[3], frr[3];
    if (i - 2 * d > 0 && i + 2 * d < ie) {
        load3(Qt, i - 2 * d, q); flux_f1(g, q, fll);
        load3(Qt, i - d, q); flux_f1(g, q, fl);
        load3(Qt, i + d, q); flux_f1(g, q, fr);
        load3(Qt, i + 2 * d, q); flux_f1(g, q, frr);
        for (int p = 0; p < 3; p++) out[p] = kD1 * (fr[p] - fl[p]) + kD2 * (frr[p] - fll[p]);
    } else if (i - 2 * d < 0) {
        load3(Qt, i - d, q); flux_f1(g, q, fl);
        load3(Qt, i, q); flux_f1(g, q, fc);
        load3(Qt, i + d, q); flux_f1(g, q, fr);
        load3(Qt, i + 2 * d, q); flux_f1(g, q, frr);
        for (int p = 0; p < 3; p++) out[p] = kD01 * fl[p] + kD11 * fc[p] + kD21 * fr[p] + kD31 * frr[p];
    } else if (i == 2 * d) {
        load3(Qt, i - 2 * d, q); flux_f1(g, q, fll);
        load3(Qt, i - d, q); flux_f1(g, q, fl);
        load3(Qt, i, q); flux_f1(g, q, fc);
        load3(Qt, i + d, q); flux_f1(g, q, fr);
        load3(Qt, i + 2 * d, q); flux_f1(g, q, frr);
        for (int p = 0; p < 3; p++) out[p] = kD02 * fll[p] + kD12 * fl[p] + kD22 * fc[p] + kD32 * fr[p] + kD2 * frr[p];
    } else if (i > ie - 2 * d) {
        // AtRightBoundary(ll, l, c, r) = AtLeftBoundary(r, c, l, ll): mirrored, not negated
        load3(Qt, i - 2 * d, q); flux_f1(g, q, fll);
        load3(Qt, i - d, q); flux_f1(g, q, fl);
        load3(Qt, i, q); flux_f1(g, q, fc);
        load3(Qt, i + d, q); flux_f1(g, q, fr);
        for (int p = 0; p < 3; p++) out[p] = kD01 * fr[p] + kD11 * fc[p] + kD21 * fl[p] + kD31 * fll[p];
    } else {
        load3(Qt, i - 2 * d, q); flux_f1(g, q, fll);
        load3(Qt, i - d, q); flux_f1(g, q, fl);
        load3(Qt, i, q); flux_f1(g, q, fc);
        load3(Qt, i + d, q); flux_f1(g, q, fr);
        load3(Qt, i + 2 * d, q); flux_f1(g, q, frr);
        for (int p = 0; p < 3; p++) out[p] = kD02 * frr[p] + kD12 * fr[p] + kD22 * fc[p] + kD32 * fl[p] + kD2 * fll[p];
    }
    return r;
}

// alpha = max |u|+c over {mask & level == max} of the current Q (main_1d_euler.cpp:86-95)
__device__ double lf_alpha(const Ctx1& c, const uint8_t* mask, const uint8_t* lvl) {
    double a = 0;
    for (int i = threadIdx.x; i < c.g.Nx; i += blockDim.x) {
        if (mask[i + G] == 0 || lvl[i] != c.g.maxlev) continue;
        double q[3];
        load3(c, c.Q, i, q);
        const double u = Div(q[0])(q[1]);
        const double cs = eos_c1(c.gamma, q);
        a = t_max(a, t_abs(u) + cs);
    }
    return block_max(a, c.red);
}

// The face flux f (between points f-1 and f) of the stage input Qt, whose six
// states f-3 .. f+2 lie in the ghosted range for every f in 0..Nx (out of
// line, as coarse_dF_at).
template <int FLUX>
__device__ __attribute__((noinline)) F3 face_at(const double* Qt, int NP, int f, double g, double alpha) {
    double s[6][3];
    for (int p = 0; p < 6; p++)
        for (int k = 0; k < 3; k++) s[p][k] = Qt[k * NP + f - 3 + p + G];
    F3 r;
    face_flux1<FLUX>(g, s, alpha, r.v);
    return r;
}

// One SSP-RK3 stage over `mask`/`lvl` (main_1d_euler.cpp:97-245).  stage 0:
// Q -= h dF; stage 1: Q = 0.75 Qc + 0.25 (Q - h dF); stage 2: Q = (Qc + 2 (Q - h dF))/3.
// FLUX: the face fluxes, WENO-Z (kRecWenoZ, the C1 configuration) or WENO5-JS
// (kRecWeno5, the shipped driver's ComputedF, WENO.h:176-184).
template <int FLUX = kRecWenoZ>
__device__ void rk_stage(const Ctx1& c, int stage, const uint8_t* mask, const uint8_t* lvl, double alpha, double dt) {
    const int Nx = c.g.Nx, NP = c.NP, maxlev = c.g.maxlev;
    for (int i = threadIdx.x; i < 3 * NP; i += blockDim.x) {
        const double v = c.Q[i];
        c.Qt[i] = v;
        if (stage == 0) c.Qc[i] = v;
    }
    __syncthreads();
    // faces f = 0..Nx (between f-1 and f) that a finest-level point of this stage needs
    for (int f = threadIdx.x; f <= Nx; f += blockDim.x) {
        const bool lw = (f >= 1) && mask[f - 1 + G] && lvl[f - 1] == maxlev;
        const bool rw = (f <= Nx - 1) && mask[f + G] && lvl[f] == maxlev;
        if (!(lw || rw)) continue;
        const F3 fl = face_at<FLUX>(c.Qt, NP, f, c.gamma, alpha);
        for (int k = 0; k < 3; k++) c.F[k * (Nx + 1) + f] = fl.v[k];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < Nx; i += blockDim.x) {
        if (mask[i + G] == 0) continue;
        const int lev = lvl[i];
        const int d = c.g.dxl(lev);
        double dF[3];
        if (lev == maxlev) {
            for (int k = 0; k < 3; k++) dF[k] = c.F[k * (Nx + 1) + i + 1] - c.F[k * (Nx + 1) + i];
        } else {
            const F3 r = coarse_dF_at(c.Qt, c.NP, c.g.Nx, c.gamma, i, d, c.flag);
            for (int k = 0; k < 3; k++) dF[k] = r.v[k];
        }
        const double h = Div(c.dx * d)(dt);
        for (int k = 0; k < 3; k++) {
            double& q = c.Q[k * NP + i + G];
            if (stage == 0) q = q - h * dF[k];
            else if (stage == 1) q = 0.75 * c.Qc[k * NP + i + G] + 0.25 * (q - h * dF[k]);
            else q = Div(3.0)(c.Qc[k * NP + i + G] + 2.0 * (q - h * dF[k]));
        }
    }
    __syncthreads();
}

// AWCM_MANAGER_1D::RemoveResidue (AWCM_MANAGER_1D.h:146-157)
__device__ void post(const Ctx1& c) {
    for (int i = threadIdx.x; i < c.g.Nx; i += blockDim.x)
        if (c.m[0][i + G] == 0)
            for (int k = 0; k < 3; k++) c.Q[k * c.NP + i + G] = 0;
    __syncthreads();
}

// Time-loop state of the instance being solved, kept in LDS for the launch.
struct LoopState {
    double t, tol, cfl, t_final;
    long long sum_active0;
    int cnt, steps, n_print;
    int bad;  // a wave speed was not finite: the state has broken down
};

// Tape writers of the adjoint recording (see Inst1D).
__device__ void tape_checkpoint(const Ctx1& c, int n) {
    Inst1D& in = *c.rec;
    if (n % in.ckpt_every) return;
    double* dst = in.ckpt + (long long)(n / in.ckpt_every) * 3 * c.NP;
    for (int i = threadIdx.x; i < 3 * c.NP; i += blockDim.x) dst[i] = c.Q[i];
}
__device__ void tape_masks(const Ctx1& c, int n) {
    uint8_t* dst = c.rec->tape_m + n * c.rec->tape_stride;
    for (int i = threadIdx.x; i < c.NP; i += blockDim.x) {
        dst[i] = c.mt[i];
        for (int k = 0; k < 4; k++) dst[(k + 1) * c.NP + i] = c.m[k][i];
    }
    for (int i = threadIdx.x; i < c.g.Nx; i += blockDim.x)
        for (int k = 0; k < 3; k++) dst[5 * c.NP + k * c.g.Nx + i] = c.lv[k][i];
    // the interpolation set PrepareEuler's InterpolateEuler just built (c.tmp)
    for (int i = threadIdx.x; i < c.NP; i += blockDim.x) dst[5 * c.NP + 3 * c.g.Nx + i] = c.tmp[i];
}

// One iteration of the driver loop (main_1d_euler.cpp:61-260).  With a tape
// attached, also records what the adjoint needs (masks, levels, dt, alphas,
// periodic checkpoints); the arithmetic is the same either way.
template <int FLUX = kRecWenoZ>
__device__ void step(const Ctx1& c, LoopState& st, const double* print_times) {
    const int rn = c.rec ? c.rec->tape_n + st.steps : 0;
    if (c.rec) tape_checkpoint(c, rn);
    prepare(c, st.tol);
    if (c.rec) tape_masks(c, rn);
    double v = 0;
    bool finite = true;
    for (int i = threadIdx.x; i < c.g.Nx; i += blockDim.x) {
        if (c.m[3][i + G] == 0) continue;
        double q[3];
        load3(c, c.Q, i, q);
        const double cs = eos_c1(c.gamma, q);
        const double u = Div(q[0])(q[1]);
        const double s = t_abs(u) + cs;
        // the reference's MAX drops a NaN depending on the scan order, so a
        // broken-down state (negative pressure, vacuum) is detected explicitly
        finite = finite && s <= 1.7976931348623157e308;
        v = t_max(v, s);
    }
    if (!finite) atomicOr(&st.bad, 1);
    const double vmax = block_max(v, c.red);  // its barriers publish st.bad
    if (st.bad || !(vmax > 0)) {
        if (threadIdx.x == 0) st.bad = 1;
        __syncthreads();
        return;
    }
    double dt = Div(vmax)(st.cfl * c.dx);
    const double t = st.t;
    const int cnt = st.cnt;
    bool do_print = false;
    if (cnt < st.n_print && t + dt > print_times[cnt]) { dt = print_times[cnt] - t; do_print = true; }
    bc(c);
    double a = lf_alpha(c, c.m[3], c.lv[3]);
    if (c.rec && threadIdx.x == 0) { c.rec->tape_s[4 * rn] = dt; c.rec->tape_s[4 * rn + 1] = a; }
    rk_stage<FLUX>(c, 0, c.m[2], c.lv[2], a, dt);
    bc(c);
    a = lf_alpha(c, c.m[2], c.lv[2]);
    if (c.rec && threadIdx.x == 0) c.rec->tape_s[4 * rn + 2] = a;
    rk_stage<FLUX>(c, 1, c.m[1], c.lv[1], a, dt);
    bc(c);
    a = lf_alpha(c, c.m[1], c.lv[1]);
    if (c.rec && threadIdx.x == 0) c.rec->tape_s[4 * rn + 3] = a;
    rk_stage<FLUX>(c, 2, c.m[0], c.lv[0], a, dt);
    post(c);
    int n0 = 0;
    for (int i = threadIdx.x; i < c.g.Nx; i += blockDim.x) n0 += c.m[0][i + G];
    __syncthreads();
    if (threadIdx.x == 0) {
        st.t = t + dt;
        if (do_print) st.cnt = cnt + 1;
        st.steps += 1;
    }
    atomicAdd((unsigned long long*)&st.sum_active0, (unsigned long long)n0);
    __syncthreads();
}

}  // namespace
}  // namespace awcm
