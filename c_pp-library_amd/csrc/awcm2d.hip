// 2-D AWCM compressible-Euler solver on gfx950: kernels and the host-side
// sequence of one step.
//
// Device layout (SoA, HBM):
//   state   4 planes (rho, rho u, rho v, E) of P*(Ny+6) doubles, P = Nx+6,
//           index (i+3) + P*(j+3) -- the reference's ghosted linear index
//           (ARRAY_2D.h:112), so every stencil address is the reference's.
//   masks   uint8, same ghosted index; mask_temp, mask0..3, a snapshot and the
//           interpolation work mask.
//   levels  uint8, the masks' ghosted index (ghost entries hold the max level
//           and are never read), so a row segment is one aligned 16-byte load.
//   lists   per stage: compacted finest-level points and coarse points,
//           packed (j << 16) | i.
//
// Kernel map (reference function -> kernels):
//   ForwardWaveletTransformationEuler (WAVELET_2D.h:478-624)   k2_lat<OP_FWT_D3/D12/C> per level
//   BandingMask + AddAdjacentZone (AWCM_MANAGER_2D.h:14-61)      k2_band_snap, k2_adjacent
//   CheckReconstruction (AWCM_MANAGER_2D.h:63-158)               k2_check per level (d3 + d1/d2 fused)
//   mask copy + ComputeLevel + AddPointForDifferentiation         k2_level_add (one pass per mask)
//     (AWCM_MANAGER_2D.h:160-241)
//   InverseWaveletTransformationEuler (WAVELET_2D.h:626-782)     k2_lat<OP_IWT_C/D12/D3> per level
//   InterpolateEuler (WAVELET_2D.h:784-907)                      k2_lat<OP_INT_D12/D3> per level
//   stage work lists                                             k2_compact
//   CFL / alpha reductions (main_2d_euler(dmach).cpp:61-104)     k2_alpha (+ dt in k2_bc)
//   ghost fill (PROBLEM_QUADRANTS.h:123-203, PROBLEM_DOUBLEMACH.h:106-245)   k2_bc
//   stage loops (main_2d_euler(dmach).cpp:106-322)               k2_stage_faces, k2_stage_update<S>
//   RK combine (:325-332)                                        folded into k2_stage_update<3>
//   RemoveResidueEuler (AWCM_MANAGER_2D.h:267-280)               k2_residue
//   optional one-workgroup tail of the coarsest levels           k2_tail<...>, k2_check_tail
// Every kernel returns at once when the step is skipped (t >= t_final), so
// many steps can be queued or graph-replayed without host round trips.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "awcm_capi.h"
#include "awcm_common.h"
#include "solver2d.h"

namespace awcm {
namespace {

constexpr int G = kGhost;
constexpr double kFluxGamma = 1.4;  // WENOZ_EULER_SOLVER_2D is default-constructed in the drivers

__device__ __forceinline__ long long gidx(const Dims2& D, int i, int j) {
    return (long long)(i + G) + (long long)D.P * (j + G);
}

__device__ __forceinline__ bool lat_point(const Lat& g, int t, int& i, int& j) {
    if (t >= g.ni * g.nj) return false;
    const int a = t % g.ni, b = t / g.ni;
    i = g.is + a * g.dx;
    j = g.js + b * g.dy;
    return true;
}

__device__ __forceinline__ void mset(const Dims2& D, uint8_t* m, int i, int j, Scal2* sc) {
    if (i < -G || i > D.Nx - 1 + G || j < -G || j > D.Ny - 1 + G) { sc->flag |= 1; return; }
    m[gidx(D, i, j)] = 1;
}

// Clamped 4-point stencil start of a detail (predict) point: WAVELET_2D.h:506-509
__device__ __forceinline__ void pstart(int i, int d, int iend_lim, int& is, int& k) {
    const int dh = d / 2;
    is = i - d - dh;
    k = 1;
    if (is < 0) is = 0, k = 0;
    else if (i + d + dh > iend_lim) is = iend_lim - 3 * d, k = 2;
}
// Clamped stencil start of a coarse (update) point: WAVELET_2D.h:595-598
__device__ __forceinline__ void ustart(int i, int d, int iend_lim, int& is, int& k) {
    const int dh = d / 2;
    is = i - d - dh;
    k = 1;
    if (is < 0) is = dh, k = 0;
    else if (i + d + dh > iend_lim) is = iend_lim - 3 * d - dh, k = 2;
}

// ------------------------------------------------------------- lattices ----
// Lattices of level lev (m = 2^(L-1-lev), h = m/2).  For every admissible size
// they coincide with the subgrid/grid ranges of MRA_GRID_UNIFORM_2D.h:108-130
// (checked against the reference formulas on the host at set-up).
enum { LAT_D1 = 0, LAT_D2 = 1, LAT_D3 = 2, LAT_C = 3 };

__device__ __forceinline__ Lat lattice(const Dims2& D, int lev, int kind) {
    const int m = 1 << (D.L - 1 - lev), h = m >> 1;
    Lat g;
    g.dx = m;
    g.dy = m;
    if (kind == LAT_D1) { g.is = h; g.ie = D.Nx - 1 - h; g.js = m; g.je = D.Ny - 1 - m; }
    else if (kind == LAT_D2) { g.is = m; g.ie = D.Nx - 1 - m; g.js = h; g.je = D.Ny - 1 - h; }
    else if (kind == LAT_D3) { g.is = h; g.ie = D.Nx - 1 - h; g.js = h; g.je = D.Ny - 1 - h; }
    else { g.is = m; g.ie = D.Nx - 1 - m; g.js = m; g.je = D.Ny - 1 - m; }
    g.ni = (g.ie - g.is) / m + 1;
    g.nj = (g.je - g.js) / m + 1;
    return g;
}

// ------------------------------------------------ lifting point operations --
// One point of one lifting pass; `act` is the point's mask byte read by the
// caller (so a thread can issue the mask loads of all its points first).
enum {
    OP_FWT_D3, OP_FWT_D12, OP_FWT_C, OP_IWT_C, OP_IWT_D12, OP_IWT_D3,
    OP_INT_D12, OP_INT_D3, OP_IAG_D12, OP_IAG_D3, OP_SINT_D12, OP_SINT_D3
};

// WAVELET_2D.h:496-523 (d3 predict + threshold)
__device__ __forceinline__ void fwt_d3_pt(const Dims2& D, int i, int j, int d, const Wts& W, double tol,
                                          double* rho, uint8_t* mask) {
    const long long p = gidx(D, i, j);
    int is, js, k1, k2;
    pstart(i, d, D.Nx - 1, is, k1);
    pstart(j, d, D.Ny - 1, js, k2);
    double v = rho[p];
    for (int c = 0; c < 4; c++) v -= rho[gidx(D, is + c * d, j)] * W.w[c][k1 + 1];
    for (int c = 0; c < 4; c++) v -= rho[gidx(D, i, js + c * d)] * W.w[c][k2 + 1];
    for (int c2 = 0; c2 < 4; c2++)
        for (int c1 = 0; c1 < 4; c1++)
            v += rho[gidx(D, is + c1 * d, js + c2 * d)] * W.w[c1][k1 + 1] * W.w[c2][k2 + 1];
    v *= 0.25;
    rho[p] = v;
    if (t_abs(v) < tol) mask[p] = 0;
}
// WAVELET_2D.h:539-583 (d1 in x / d2 in y predict + threshold)
__device__ __forceinline__ void fwt_d12_pt(const Dims2& D, int kind, int i, int j, int d, const Wts& W, double tol,
                                           double* rho, uint8_t* mask) {
    const long long p = gidx(D, i, j);
    double v = rho[p];
    int s, k;
    if (kind == LAT_D1) {
        pstart(i, d, D.Nx - 1, s, k);
        for (int c = 0; c < 4; c++) v -= rho[gidx(D, s + c * d, j)] * W.w[c][k + 1];
    } else {
        pstart(j, d, D.Ny - 1, s, k);
        for (int c = 0; c < 4; c++) v -= rho[gidx(D, i, s + c * d)] * W.w[c][k + 1];
    }
    v *= 0.5;
    rho[p] = v;
    if (t_abs(v) < tol) mask[p] = 0;
}
// WAVELET_2D.h:584-622 (update, +) and :644-683 (undo, -)
template <int SIGN>
__device__ __forceinline__ void lift_c_pt(const Dims2& D, int i, int j, int d, const Wts& W, double* rho) {
    const long long p = gidx(D, i, j);
    int is, js, k1, k2;
    ustart(i, d, D.Nx - 1, is, k1);
    ustart(j, d, D.Ny - 1, js, k2);
    double v = rho[p];
    for (int c = 0; c < 4; c++) {
        const double t = rho[gidx(D, is + c * d, j)] * W.w[c][k1 + 1];
        if (SIGN > 0) v += t; else v -= t;
    }
    for (int c = 0; c < 4; c++) {
        const double t = rho[gidx(D, i, js + c * d)] * W.w[c][k2 + 1];
        if (SIGN > 0) v += t; else v -= t;
    }
    for (int c2 = 0; c2 < 4; c2++)
        for (int c1 = 0; c1 < 4; c1++) {
            const double t = rho[gidx(D, is + c1 * d, js + c2 * d)] * W.w[c1][k1 + 1] * W.w[c2][k2 + 1];
            if (SIGN > 0) v += t; else v -= t;
        }
    rho[p] = v;
}
// WAVELET_2D.h:688-734 (inverse d2 / d1)
__device__ __forceinline__ void iwt_d12_pt(const Dims2& D, int kind, int i, int j, int d, const Wts& W, double* rho) {
    const long long p = gidx(D, i, j);
    double v = rho[p];
    v *= 2.0;
    int s, k;
    if (kind == LAT_D1) {
        pstart(i, d, D.Nx - 1, s, k);
        for (int c = 0; c < 4; c++) v += rho[gidx(D, s + c * d, j)] * W.w[c][k + 1];
    } else {
        pstart(j, d, D.Ny - 1, s, k);
        for (int c = 0; c < 4; c++) v += rho[gidx(D, i, s + c * d)] * W.w[c][k + 1];
    }
    rho[p] = v;
}
// WAVELET_2D.h:738-778 (inverse d3)
__device__ __forceinline__ void iwt_d3_pt(const Dims2& D, int i, int j, int d, const Wts& W, double* rho) {
    const long long p = gidx(D, i, j);
    int is, js, k1, k2;
    pstart(i, d, D.Nx - 1, is, k1);
    pstart(j, d, D.Ny - 1, js, k2);
    double v = rho[p];
    v *= 4.0;
    for (int c = 0; c < 4; c++) v += rho[gidx(D, is + c * d, j)] * W.w[c][k1 + 1];
    for (int c = 0; c < 4; c++) v += rho[gidx(D, i, js + c * d)] * W.w[c][k2 + 1];
    for (int c2 = 0; c2 < 4; c2++)
        for (int c1 = 0; c1 < 4; c1++)
            v -= rho[gidx(D, is + c1 * d, js + c2 * d)] * W.w[c1][k1 + 1] * W.w[c2][k2 + 1];
    rho[p] = v;
}
// WAVELET_2D.h:808-905 (interpolation of momentum and energy, d1/d2 and d3:
// components C0..C1-1 = 1..3), :937-1039 (InterpolateAllGridEuler: all four)
// and :372-470 (Interpolate of the scalar field: component 0).  All stencil
// values of the components are loaded before any store, so the loads of one
// component do not wait behind the store of another.
template <int C0, int C1>
__device__ __forceinline__ void int_d12_pt(const Dims2& D, int kind, int i, int j, int d, const Wts& W, double* Q) {
    const long long p = gidx(D, i, j);
    int s, k;
    if (kind == LAT_D1) pstart(i, d, D.Nx - 1, s, k);
    else pstart(j, d, D.Ny - 1, s, k);
    double x[C1 - C0][4];
    for (int c4 = C0; c4 < C1; c4++) {
        const double* q = Q + c4 * D.plane;
        for (int c = 0; c < 4; c++)
            x[c4 - C0][c] = (kind == LAT_D1) ? q[gidx(D, s + c * d, j)] : q[gidx(D, i, s + c * d)];
    }
    for (int c4 = C0; c4 < C1; c4++) {
        double v = 0.0;
        for (int c = 0; c < 4; c++) v += x[c4 - C0][c] * W.w[c][k + 1];
        Q[p + c4 * D.plane] = v;
    }
}
template <int C0, int C1>
__device__ __forceinline__ void int_d3_pt(const Dims2& D, int i, int j, int d, const Wts& W, double* Q) {
    const long long p = gidx(D, i, j);
    int is, js, k1, k2;
    pstart(i, d, D.Nx - 1, is, k1);
    pstart(j, d, D.Ny - 1, js, k2);
    double row[C1 - C0][4], col[C1 - C0][4], blk[C1 - C0][16];
    for (int c4 = C0; c4 < C1; c4++) {
        const double* q = Q + c4 * D.plane;
        for (int c = 0; c < 4; c++) row[c4 - C0][c] = q[gidx(D, is + c * d, j)];
        for (int c = 0; c < 4; c++) col[c4 - C0][c] = q[gidx(D, i, js + c * d)];
        for (int c2 = 0; c2 < 4; c2++)
            for (int c1 = 0; c1 < 4; c1++) blk[c4 - C0][4 * c2 + c1] = q[gidx(D, is + c1 * d, js + c2 * d)];
    }
    for (int c4 = C0; c4 < C1; c4++) {
        double v = 0.0;
        for (int c = 0; c < 4; c++) v += row[c4 - C0][c] * W.w[c][k1 + 1];
        for (int c = 0; c < 4; c++) v += col[c4 - C0][c] * W.w[c][k2 + 1];
        for (int c2 = 0; c2 < 4; c2++)
            for (int c1 = 0; c1 < 4; c1++) v -= blk[c4 - C0][4 * c2 + c1] * W.w[c1][k1 + 1] * W.w[c2][k2 + 1];
        Q[p + c4 * D.plane] = v;
    }
}

// Dispatch of one lattice point of pass OP.  For the interpolation passes of
// PrepareEuler `mask` is the work mask (mask3) and `aux` the old mask
// (mask_temp); for those of Interpolate / InterpolateEuler(from, to) `mask`
// is the work mask (CheckReconstruction of to minus from) and `aux` is from.  For those of InterpolateAllGridEuler `aux` is mask_from: the
// reference's work mask is CheckReconstruction of (every point but mask_from's
// finest-grid points), tested as `mask_temp && !mask_from`; the closure only
// adds points, and the points it can add are all in mask_from, so the test is
// `!mask_from` alone (WAVELET_2D.h:917-922,943).
//
// For the inverse transform of PrepareEuler `aux` is mask_temp (the previous
// mask0): a point inactive in mask3 gets rho = 0, which only needs a store
// where rho can be non-zero -- at points of the previous mask0, or anywhere
// when the host wrote Q or mask0 since the last residue pass (`full`, the
// step's residue_full) -- because between steps Q is zero outside mask0 and
// the forward transform writes only active points.  Code 2 marks a point to
// zero, 1 a point to transform.  Without `aux` (the scalar Prepare, whose
// mask_temp is not the previous mask0) every inactive point is zeroed.
template <int OP>
__device__ __forceinline__ uint8_t op_active(const uint8_t* mask, const uint8_t* aux, long long p, int full) {
    if (OP == OP_INT_D12 || OP == OP_INT_D3 || OP == OP_SINT_D12 || OP == OP_SINT_D3)
        return (mask[p] != 0 && aux[p] == 0) ? 1 : 0;
    if (OP == OP_IAG_D12 || OP == OP_IAG_D3) return aux[p] == 0 ? 1 : 0;
    if (OP == OP_IWT_C || OP == OP_IWT_D12 || OP == OP_IWT_D3) {
        const uint8_t a = mask[p], was = aux ? aux[p] : 1;  // both loads independent
        return a ? 1 : (full || was) ? 2 : 0;
    }
    return mask[p];
}
template <int OP>
__device__ __forceinline__ void op_point(const Dims2& D, int kind, int i, int j, int d, uint8_t act, const Wts& W,
                                         double tol, double* Q, uint8_t* mask) {
    if (OP == OP_IWT_C || OP == OP_IWT_D12 || OP == OP_IWT_D3) {
        if (act == 2) Q[gidx(D, i, j)] = 0;
        if (act != 1) return;
    } else if (!act) {
        return;
    }
    if (OP == OP_FWT_D3) fwt_d3_pt(D, i, j, d, W, tol, Q, mask);
    else if (OP == OP_FWT_D12) fwt_d12_pt(D, kind, i, j, d, W, tol, Q, mask);
    else if (OP == OP_FWT_C) lift_c_pt<1>(D, i, j, d, W, Q);
    else if (OP == OP_IWT_C) lift_c_pt<-1>(D, i, j, d, W, Q);
    else if (OP == OP_IWT_D12) iwt_d12_pt(D, kind, i, j, d, W, Q);
    else if (OP == OP_IWT_D3) iwt_d3_pt(D, i, j, d, W, Q);
    else if (OP == OP_INT_D12) int_d12_pt<1, 4>(D, kind, i, j, d, W, Q);
    else if (OP == OP_INT_D3) int_d3_pt<1, 4>(D, i, j, d, W, Q);
    else if (OP == OP_IAG_D12) int_d12_pt<0, 4>(D, kind, i, j, d, W, Q);
    else if (OP == OP_IAG_D3) int_d3_pt<0, 4>(D, i, j, d, W, Q);
    else if (OP == OP_SINT_D12) int_d12_pt<0, 1>(D, kind, i, j, d, W, Q);
    else int_d3_pt<0, 1>(D, i, j, d, W, Q);
}
template <int OP>
__device__ __forceinline__ int op_kind1() {
    return (OP == OP_FWT_D3 || OP == OP_IWT_D3 || OP == OP_INT_D3 || OP == OP_IAG_D3 || OP == OP_SINT_D3) ? LAT_D3
           : (OP == OP_FWT_C || OP == OP_IWT_C)                     ? LAT_C
                                                                     : LAT_D1;
}
template <int OP>
__device__ __forceinline__ bool op_two() {
    return OP == OP_FWT_D12 || OP == OP_IWT_D12 || OP == OP_INT_D12 || OP == OP_IAG_D12 || OP == OP_SINT_D12;
}

// One lifting pass of one level: each thread takes K consecutive lattice points
// of a row, reads their K mask bytes first, then works on the active ones.
template <int OP, int K>
__global__ void __launch_bounds__(256) k2_lat(Dims2 D, int lev, Wts W, double tol, double* __restrict__ Q,
                                              uint8_t* __restrict__ mask, const uint8_t* __restrict__ aux,
                                              const Scal2* sc) {
    const int skip = sc->skip;  // tested after the mask loads
    const int full = sc->residue_full;
    const int d = 1 << (D.L - 1 - lev);
    int kind = op_kind1<OP>();
    Lat g = lattice(D, lev, kind);
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (op_two<OP>()) {
        const int n1 = ((g.ni + K - 1) / K) * g.nj;
        if (t >= n1) { t -= n1; kind = LAT_D2; g = lattice(D, lev, LAT_D2); }
    }
    const int segs = (g.ni + K - 1) / K;
    if (t >= segs * g.nj) return;
    const int b = t / segs, a = t % segs;
    const int j = g.js + b * g.dy;
    uint8_t act[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int ii = a * K + k;
        act[k] = ii < g.ni ? op_active<OP>(mask, aux, gidx(D, g.is + ii * g.dx, j), full) : 0;
    }
    if (skip) return;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int ii = a * K + k;
        if (ii < g.ni) op_point<OP>(D, kind, g.is + ii * g.dx, j, d, act[k], W, tol, Q, mask);
    }
}

// The same pass over every point of one lattice, by one workgroup (tail levels).
template <int OP>
__device__ void wg_pass(const Dims2& D, int lev, const Wts& W, double tol, double* Q, uint8_t* mask, const uint8_t* aux) {
    const int d = 1 << (D.L - 1 - lev);
    const int nk = op_two<OP>() ? 2 : 1;
    for (int q = 0; q < nk; q++) {
        const int kind = q ? LAT_D2 : op_kind1<OP>();
        const Lat g = lattice(D, lev, kind);
        const int n = g.ni * g.nj;
        constexpr int B = 4;  // points per thread per round, masks loaded first
        for (int t0 = threadIdx.x * B; t0 < n; t0 += blockDim.x * B) {
            uint8_t act[B];
#pragma unroll
            for (int k = 0; k < B; k++) {
                const int t = t0 + k;
                act[k] = t < n ? op_active<OP>(mask, aux, gidx(D, g.is + (t % g.ni) * g.dx, g.js + (t / g.ni) * g.dy), 1) : 0;
            }
#pragma unroll
            for (int k = 0; k < B; k++) {
                const int t = t0 + k;
                if (t < n) op_point<OP>(D, kind, g.is + (t % g.ni) * g.dx, g.js + (t / g.ni) * g.dy, d, act[k], W, tol, Q, mask);
            }
        }
    }
}

// Tail of a level cascade: levels 0..T of the forward transform (finest of
// them first), the inverse transform or the interpolation (coarsest first),
// in one workgroup with a barrier between passes.
enum { TAIL_FWT = 0, TAIL_IWT = 1, TAIL_INT = 2 };
template <int WHICH>
__global__ void __launch_bounds__(1024) k2_tail(Dims2 D, int T, Wts W, double tol, double* __restrict__ Q,
                                                uint8_t* __restrict__ mask, const uint8_t* __restrict__ aux,
                                                const Scal2* sc) {
    if (sc->skip) return;
    if (WHICH == TAIL_FWT) {
        for (int lev = T; lev >= 0; lev--) {
            wg_pass<OP_FWT_D3>(D, lev, W, tol, Q, mask, aux);
            __syncthreads();
            wg_pass<OP_FWT_D12>(D, lev, W, tol, Q, mask, aux);
            __syncthreads();
            wg_pass<OP_FWT_C>(D, lev, W, tol, Q, mask, aux);
            __syncthreads();
        }
    } else if (WHICH == TAIL_IWT) {
        for (int lev = 0; lev <= T; lev++) {
            wg_pass<OP_IWT_C>(D, lev, W, tol, Q, mask, aux);
            __syncthreads();
            wg_pass<OP_IWT_D12>(D, lev, W, tol, Q, mask, aux);
            __syncthreads();
            wg_pass<OP_IWT_D3>(D, lev, W, tol, Q, mask, aux);
            __syncthreads();
        }
    } else {
        for (int lev = 0; lev <= T; lev++) {
            wg_pass<OP_INT_D12>(D, lev, W, tol, Q, mask, aux);
            __syncthreads();
            wg_pass<OP_INT_D3>(D, lev, W, tol, Q, mask, aux);
            __syncthreads();
        }
    }
}

// ----------------------------------------- reconstruction check (fused) ----
// AWCM_MANAGER_2D::CheckReconstruction (AWCM_MANAGER_2D.h:63-158), one level.
// A cell (I, J) of level lev holds the d3 point (h+I m, h+J m), the d1 point
// (h+I m, (J+1) m) and the d2 point ((I+1) m, h+J m).  The reference runs the
// d3 pass before the d1/d2 passes because a d3 stencil activates d1/d2 points
// of the same level; here each d1/d2 point instead gathers that activation
// from the (at most six) d3 points whose clamped stencil reaches it, so all
// three sub-bands of a level run in one pass.  Every write is a 1 and d3
// points are neither written nor read across levels within the pass.
__device__ __forceinline__ int clamp_start(int c, int d, int lim) {
    int s = c - d - d / 2;
    if (s < 0) s = 0;
    else if (c + d + d / 2 > lim) s = lim - 3 * d;
    return s;
}

// every mask byte a cell needs, loaded up front (independent loads)
struct CellMasks {
    uint8_t a3, o1, o2, c1[6], c2[6];
};
__device__ __forceinline__ CellMasks check_cell_load(const Dims2& D, int lev, int I, int J, const uint8_t* mask) {
    const int m = 1 << (D.L - 1 - lev), h = m >> 1;
    const int nI = (D.Nx - 1) / m, nJ = (D.Ny - 1) / m;
    const int i3 = h + I * m, j3 = h + J * m;
    const bool has1 = J + 1 <= nJ - 1, has2 = I + 1 <= nI - 1;
    const int j1 = (J + 1) * m, i2 = (I + 1) * m;
    CellMasks c;
    c.a3 = mask[gidx(D, i3, j3)];
    c.o1 = has1 ? mask[gidx(D, i3, j1)] : 0;
    c.o2 = has2 ? mask[gidx(D, i2, j3)] : 0;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int o = 2 * q - 5;
        const int jj = j1 + o * h, ii = i2 + o * h;
        c.c1[q] = (has1 && jj >= h && jj <= D.Ny - 1 - h) ? mask[gidx(D, i3, jj)] : 0;
        c.c2[q] = (has2 && ii >= h && ii <= D.Nx - 1 - h) ? mask[gidx(D, ii, j3)] : 0;
    }
    return c;
}

__device__ __forceinline__ void check_cell_apply(const Dims2& D, int lev, int I, int J, const CellMasks& cm,
                                                 uint8_t* mask, Scal2* sc) {
    const int m = 1 << (D.L - 1 - lev), h = m >> 1;
    const int i3 = h + I * m, j3 = h + J * m;
    const int j1 = (J + 1) * m, i2 = (I + 1) * m;
    const uint8_t a3 = cm.a3, o1 = cm.o1, o2 = cm.o2;
    const uint8_t* c1 = cm.c1;
    const uint8_t* c2 = cm.c2;
    // d1 / d2 points activated by the clamped stencil of an active d3 point
    bool a1 = o1 != 0, a2 = o2 != 0;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int o = 2 * q - 5;
        if (c1[q]) {
            const int s = clamp_start(j1 + o * h, m, D.Ny - 1);
            if (j1 >= s && j1 <= s + 3 * m) a1 = true;
        }
        if (c2[q]) {
            const int s = clamp_start(i2 + o * h, m, D.Nx - 1);
            if (i2 >= s && i2 <= s + 3 * m) a2 = true;
        }
    }
    if (a3) {
        const int is = clamp_start(i3, m, D.Nx - 1), js = clamp_start(j3, m, D.Ny - 1);
        for (int c = 0; c < 4; c++) mset(D, mask, is + c * m, j3, sc);
        for (int c = 0; c < 4; c++) mset(D, mask, i3, js + c * m, sc);
        for (int cc2 = 0; cc2 < 4; cc2++)
            for (int cc1 = 0; cc1 < 4; cc1++) mset(D, mask, is + cc1 * m, js + cc2 * m, sc);
    }
    if (a1) {
        const int is = clamp_start(i3, m, D.Nx - 1);
        for (int c = 0; c < 4; c++) mset(D, mask, is + c * m, j1, sc);
    }
    if (a2) {
        const int js = clamp_start(j3, m, D.Ny - 1);
        for (int c = 0; c < 4; c++) mset(D, mask, i2, js + c * m, sc);
    }
}

__device__ __forceinline__ void check_cell(const Dims2& D, int lev, int I, int J, uint8_t* mask, Scal2* sc) {
    check_cell_apply(D, lev, I, J, check_cell_load(D, lev, I, J, mask), mask, sc);
}

// The skip flag is read beside the mask loads, not before them: a skipped
// step costs the loads, a live one saves a round trip.
template <int K>
__global__ void __launch_bounds__(256) k2_check(Dims2 D, int lev, uint8_t* __restrict__ mask, Scal2* sc) {
    const int skip = sc->skip;
    const int m = 1 << (D.L - 1 - lev);
    const int nI = (D.Nx - 1) / m, nJ = (D.Ny - 1) / m;
    const int segs = (nI + K - 1) / K;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= segs * nJ) return;
    const int J = t / segs, I0 = (t % segs) * K;
    CellMasks cm[K];
#pragma unroll
    for (int k = 0; k < K; k++)
        if (I0 + k < nI) cm[k] = check_cell_load(D, lev, I0 + k, J, mask);
    if (skip) return;
#pragma unroll
    for (int k = 0; k < K; k++)
        if (I0 + k < nI) check_cell_apply(D, lev, I0 + k, J, cm[k], mask, sc);
}

// Levels T..0 of the check in one workgroup.
__global__ void __launch_bounds__(1024) k2_check_tail(Dims2 D, int T, uint8_t* __restrict__ mask, Scal2* sc) {
    if (sc->skip) return;
    for (int lev = T; lev >= 0; lev--) {
        const int m = 1 << (D.L - 1 - lev);
        const int nI = (D.Nx - 1) / m, nJ = (D.Ny - 1) / m;
        for (int t = threadIdx.x; t < nI * nJ; t += blockDim.x) check_cell(D, lev, t % nI, t / nI, mask, sc);
        __syncthreads();
    }
}

// ------------------------------------------------------ dense mask passes --
// Dense passes take 16 consecutive points of a row per thread: device rows are
// 16-byte aligned (pitch P is a multiple of 16), so each mask row segment is
// one 16-byte load.  Thread t: row jr = t / (P/16) (ghost rows included),
// segment c: ghosted columns 16c .. 16c+15 (i = 16c - 3 + k).
struct Seg {
    long long p0;  // linear index of the segment's first byte
    int i0, j;     // i of the first byte, row j
};
__device__ __forceinline__ bool seg_of(const Dims2& D, long long t, Seg& s) {
    const int spr = D.P / 16;
    const long long nseg = (long long)spr * (D.Ny + 2 * G);
    if (t >= nseg) return false;
    const int jr = int(t / spr), c = int(t % spr);
    s.p0 = (long long)jr * D.P + 16 * c;
    s.i0 = 16 * c - G;
    s.j = jr - G;
    return true;
}
union V16 {
    uint4 v;
    uint8_t b[16];
};
__device__ __forceinline__ V16 ld16(const uint8_t* p) { V16 r; r.v = *reinterpret_cast<const uint4*>(p); return r; }
__device__ __forceinline__ void st16(uint8_t* p, const V16& x) { *reinterpret_cast<uint4*>(p) = x.v; }

// mask_temp = mask0 (whole ghosted buffer; dst == src when mask_temp is to be
// left alone), masks 1..3 cleared (z1 non-null) and the step's scalar resets.
// The residue pass of a step (or of PostProcessEuler) zeroes only points that
// the step's masks touch, unless `dirty` says Q may be non-zero elsewhere; a
// PrepareEuler on its own (prepare_only) runs no residue pass but drops points
// from mask0, so it leaves `dirty` set for the next pass.
// Every thread decides the skip itself, from t and t_final (which only earlier
// launches write): a skipped step leaves mask_temp and masks 1..3 as the last
// step left them.
__global__ void k2_begin_copy(Dims2 D, const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, double* alpha,
                              uint8_t* z1, uint8_t* z2, uint8_t* z3, Scal2* sc, int force, int prepare_only) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int skip = (!force && !(sc->t < sc->t_final)) ? 1 : 0;
    if (t == 0) {
        sc->skip = skip;
        if (!skip) { sc->residue_full = sc->dirty; sc->dirty = prepare_only ? 1 : 0; }
        sc->vmax = 0ull;
        for (int k = 0; k < 3; k++) { sc->cntW[k] = 0; sc->cntC[k] = 0; sc->cntXS[k] = 0; sc->cntYS[k] = 0; }
    }
    if (skip) return;
    const int na = 3 * (D.Nx + D.Ny);
    if (t < na) alpha[t] = 0.0;
    Seg s;
    if (seg_of(D, t, s)) {
        st16(dst + s.p0, ld16(src + s.p0));
        // masks 1..3 are rebuilt from mask0 this step; clearing them here lets
        // the fused copy/level/add-point pass write nothing but ones
        if (z1) {
            V16 zero;
            zero.v = make_uint4(0, 0, 0, 0);
            st16(z1 + s.p0, zero);
            st16(z2 + s.p0, zero);
            st16(z3 + s.p0, zero);
        }
    }
}

// BandingMask (2-wide band) on mask0, then the AddAdjacentZone snapshot
__global__ void k2_band_snap(Dims2 D, uint8_t* __restrict__ m0, uint8_t* __restrict__ snap, const Scal2* sc) {
    const int skip = sc->skip;  // tested after the mask load
    Seg s;
    if (!seg_of(D, (long long)blockIdx.x * blockDim.x + threadIdx.x, s)) return;
    V16 v = ld16(m0 + s.p0);
    if (skip) return;
    const bool row_band = s.j >= 0 && s.j <= D.Ny - 1 && (s.j < 2 || s.j > D.Ny - 1 - 2);
    bool changed = false;
    if (s.j >= 0 && s.j <= D.Ny - 1) {
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int i = s.i0 + k;
            if (i >= 0 && i <= D.Nx - 1 && (row_band || i < 2 || i > D.Nx - 1 - 2) && !v.b[k]) { v.b[k] = 1; changed = true; }
        }
    }
    if (changed) st16(m0 + s.p0, v);
    st16(snap + s.p0, v);
}

// work = to, cleared at the finest-grid points of from (WAVELET_2D.h:364-368):
// the start of the work mask of Interpolate / InterpolateEuler(from, to)
__global__ void k2_to_minus_from(Dims2 D, const uint8_t* __restrict__ from, const uint8_t* __restrict__ to,
                                 uint8_t* __restrict__ work) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.plane) return;
    const int i = int(t % D.P) - G, j = int(t / D.P) - G;
    const bool fine = i >= 0 && i <= D.Nx - 1 && j >= 0 && j <= D.Ny - 1;
    work[t] = (fine && from[t]) ? 0 : to[t];
}

// Sub-band of an interior point: level lev (0..L-2) and mu (1: odd i, 2: odd j,
// 3: both), or -1 for coarsest-grid and boundary points.  Equivalent to
// membership of the subgrid(lev, mu) lattices of MRA_GRID_UNIFORM_2D.h:114-130.
__device__ __forceinline__ int classify(int i, int j, const Dims2& D, int& mu) {
    if (i <= 0 || j <= 0 || i >= D.Nx - 1 || j >= D.Ny - 1) return -1;
    const int a = __builtin_ctz(i), b = __builtin_ctz(j);
    const int s = a < b ? a : b;
    const int lev = D.L - 2 - s;
    if (lev < 0) return -1;
    mu = (a == s && b == s) ? 3 : (a == s ? 1 : 2);
    return lev;
}

// Row segments of B bytes (B = 2, 4, 8, 16) of the ghosted layout; P is a
// multiple of 16, so segments never straddle rows.
template <int B> struct VecB;
template <> struct VecB<2> { using T = uint16_t; };
template <> struct VecB<4> { using T = uint32_t; };
template <> struct VecB<8> { using T = uint2; };
template <> struct VecB<16> { using T = uint4; };
template <int B>
union UB {
    typename VecB<B>::T v;
    uint8_t b[B];
};
template <int B>
__device__ __forceinline__ UB<B> ldb(const uint8_t* p) {
    UB<B> r;
    r.v = *reinterpret_cast<const typename VecB<B>::T*>(p);
    return r;
}
template <int B>
__device__ __forceinline__ void stb(uint8_t* p, const UB<B>& x) {
    *reinterpret_cast<typename VecB<B>::T*>(p) = x.v;
}
template <int B>
__device__ __forceinline__ bool segb_of(const Dims2& D, long long t, Seg& s) {
    const int spr = D.P / B;
    const long long nseg = (long long)spr * (D.Ny + 2 * G);
    if (t >= nseg) return false;
    const int jr = int(t / spr), c = int(t % spr);
    s.p0 = (long long)jr * D.P + B * c;
    s.i0 = B * c - G;
    s.j = jr - G;
    return true;
}
// bit k set where byte k of the segment is nonzero
template <int B>
__device__ __forceinline__ uint32_t nonzero_bits(const UB<B>& v) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < B; k++) m |= v.b[k] ? 1u << k : 0u;
    return m;
}

// Bytes per thread of the adjacent-zone and level/add-point passes: a thread
// works through its segment's active points one after another, so wide
// segments serialise dense regions, narrow ones pay per-thread overhead on
// the (mostly empty) rest of the grid.
constexpr int kMaskB = 4;

// AWCM_MANAGER_2D::AddAdjacentZone (AWCM_MANAGER_2D.h:14-61), every level at
// once from the snapshot, a B-point row segment per thread.
template <int B>
__global__ void k2_adjacent(Dims2 D, const uint8_t* __restrict__ snap, uint8_t* __restrict__ mask, Scal2* sc) {
    const int skip = sc->skip;  // tested after the snapshot load
    Seg s;
    if (!segb_of<B>(D, (long long)blockIdx.x * blockDim.x + threadIdx.x, s)) return;
    if (s.j < 1 || s.j > D.Ny - 2) return;  // no interior point (classify is -1)
    uint32_t vm = nonzero_bits<B>(ldb<B>(snap + s.p0));
    if (skip) return;
    const int j = s.j;
    for (; vm; vm &= vm - 1) {
        const int i = s.i0 + __builtin_ctz(vm);
        int mu;
        const int lev = classify(i, j, D, mu);  // -1 for ghost, padding and boundary bytes
        if (lev < 0) continue;
        if (lev == D.maxlev - 1) {
            mset(D, mask, i - 1, j, sc); mset(D, mask, i + 1, j, sc);
            mset(D, mask, i, j - 1, sc); mset(D, mask, i, j + 1, sc);
        } else {
            const int m = 1 << (D.L - 1 - lev);
            const int dq = m / 4, dh = m / 2;
            mset(D, mask, i - dq, j, sc); mset(D, mask, i + dq, j, sc);
            mset(D, mask, i, j - dq, sc); mset(D, mask, i, j + dq, sc);
            mset(D, mask, i - dh, j, sc); mset(D, mask, i + dh, j, sc);
            mset(D, mask, i, j - dh, sc); mset(D, mask, i, j + dh, sc);
        }
    }
}

// One pass per mask: mask_{k+1} = mask_k (ghosts included), ComputeLevel(level_k,
// mask_k) (AWCM_MANAGER_2D.h:215-241) and AddPointForDifferentiation(mask_{k+1},
// level_k) (AWCM_MANAGER_2D.h:160-213).  mask_{k+1} was cleared at the start of
// the step, so the copy and the scatter both write only ones and cannot undo
// each other.  The reference snapshots mask_{k+1} right after copying it from
// mask_k, so mask_k is the snapshot; a point's own level is all the scatter
// needs.  A point with no active neighbour keeps its previous level, as in the
// reference.
// A thread takes a B-point row segment of the ghosted rows and is the only
// writer of its segment of level_k (mask_{k+1} is written with single-byte
// stores of ones, which other segments' scatters may also write).  The
// nearest-neighbour test of the segment's points comes from the segment, the
// rows above and below and the two edge bytes; only isolated points search
// wider.
template <int B>
__global__ void k2_level_add(Dims2 D, const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                             uint8_t* __restrict__ lvl, Scal2* sc) {
    constexpr uint32_t kAll = (1u << B) - 1u;
    const int skip = sc->skip;  // tested after the mask load
    Seg s;
    if (!segb_of<B>(D, (long long)blockIdx.x * blockDim.x + threadIdx.x, s)) return;
    const UB<B> v = ldb<B>(src + s.p0);
    if (skip) return;
    // bytes past the last ghost column are row padding, not mask points
    const int last = D.Nx - 1 + G - s.i0;
    const uint32_t vm = nonzero_bits<B>(v) & (last < 0 ? 0u : last >= B - 1 ? kAll : (2u << last) - 1u);
    const bool row_in = s.j >= 0 && s.j <= D.Ny - 1;
    // interior columns of the segment: k in [kb, ke]
    const int kb = s.i0 < 0 ? -s.i0 : 0;
    const int ke = D.Nx - 1 - s.i0 < B - 1 ? D.Nx - 1 - s.i0 : B - 1;
    const uint32_t im = (row_in && ke >= kb) ? (((2u << ke) - 1u) & ~((1u << kb) - 1u)) : 0u;
    for (uint32_t w = vm & ~im; w; w &= w - 1) dst[s.p0 + __builtin_ctz(w)] = 1;  // ghost points: copy only
    if (!im) return;
    const uint32_t act = vm & im;
    if (!act) {
        if (im == kAll) {
            UB<B> z;
#pragma unroll
            for (int k = 0; k < B; k++) z.b[k] = 0;
            stb<B>(lvl + s.p0, z);
        } else {
            for (uint32_t w = im; w; w &= w - 1) lvl[s.p0 + __builtin_ctz(w)] = 0;
        }
        return;
    }
    UB<B> L = ldb<B>(lvl + s.p0);
    const UB<B> up = ldb<B>(src + s.p0 - D.P), dn = ldb<B>(src + s.p0 + D.P);
    // edge bytes: live whenever the segment's first or last point is interior
    const uint8_t lft = (act & 1u) ? src[s.p0 - 1] : 0;
    const uint8_t rgt = (act >> (B - 1)) & 1u ? src[s.p0 + B] : 0;
    uint32_t nbm = 0;
#pragma unroll
    for (int k = 0; k < B; k++) {
        const uint8_t l = k > 0 ? v.b[k - 1] : lft, r = k < B - 1 ? v.b[k + 1] : rgt;
        nbm |= (l | r | up.b[k] | dn.b[k]) ? 1u << k : 0u;
        if (((im & ~vm) >> k) & 1u) L.b[k] = 0;  // inactive interior point
    }
    if (act == kAll && s.i0 >= 1 && s.i0 + B - 1 <= D.Nx - 2 && s.j >= 1 && s.j <= D.Ny - 2) {
        // a run of B active interior points: each has an active neighbour in
        // the run, so each is at the finest level, and the union of their +-3
        // crosses is the row span i0-3 .. i0+B+2 plus the segment's own
        // columns in rows j-3 .. j+3 -- aligned B-byte stores of ones where
        // the loop below would store byte by byte
        UB<B> top, one;
#pragma unroll
        for (int k = 0; k < B; k++) { top.b[k] = uint8_t(D.maxlev); one.b[k] = 1; }
        stb<B>(lvl + s.p0, top);
#pragma unroll
        for (int c = -3; c <= 3; c++) stb<B>(dst + s.p0 + (long long)c * D.P, one);
#pragma unroll
        for (int c = 1; c <= 3; c++) { dst[s.p0 - c] = 1; dst[s.p0 + B - 1 + c] = 1; }
        return;
    }
    stb<B>(lvl + s.p0, L);  // active points are rewritten below, in program order
    const int j = s.j;
    const int NX = D.Nx - 1, NY = D.Ny - 1;
    for (uint32_t w = act; w; w &= w - 1) {
        const int k = __builtin_ctz(w);
        const long long t = s.p0 + k;
        const int i = s.i0 + k;
        dst[t] = 1;
        int lev;
        if ((nbm >> k) & 1u) {
            lev = D.maxlev;  // an active neighbour at distance 1 (always inside the ghosted grid)
        } else {
            lev = lvl[t];  // kept when no neighbour is active
            int d = 1;
            for (int l = D.maxlev; l >= 0; l--) {
                if (i - d < -G || i + d > D.Nx - 1 + G || j - d < -G || j + d > D.Ny - 1 + G) { sc->flag |= 1; break; }
                if (src[t - d] || src[t + d] || src[t - (long long)d * D.P] || src[t + (long long)d * D.P]) { lev = l; break; }
                d *= 2;
            }
        }
        lvl[t] = (uint8_t)lev;
        if (i < 1 || i > D.Nx - 2 || j < 1 || j > D.Ny - 2) continue;
        const int dd = 1 << (D.L - 1 - lev);  // grid[lev].dx
        if (lev == D.maxlev) {
            for (int c = -3; c <= 3; c++) { mset(D, dst, i + c, j, sc); mset(D, dst, i, j + c, sc); }
            continue;
        }
        int lo, hi;
        if (i - 2 * dd < 0) lo = -1, hi = 2;
        else if (i + 2 * dd > NX) lo = -2, hi = 1;
        else lo = -2, hi = 2;
        for (int c = lo; c <= hi; c++) mset(D, dst, i + dd * c, j, sc);
        if (j - 2 * dd < 0) lo = -1, hi = 2;
        else if (j + 2 * dd > NY) lo = -2, hi = 1;
        else lo = -2, hi = 2;
        for (int c = lo; c <= hi; c++) mset(D, dst, i, j + dd * c, sc);
    }
}

// --------------------------------------------------------- compaction ------
// Stage work lists, for each stage mask k = 0..2:
//   W  points of mask_k at the finest level (WENO-Z faces),
//   C  points of mask_k at coarser levels (central differences),
//   XS W points whose left neighbour is not a W point (their left x-face is
//      not the right face of another W point),
//   YS W points whose bottom neighbour is not a W point.
// A thread takes a 16-point row segment, so entries come out in row order
// within a thread; one atomic per list per block.  List order does not affect
// any result.
constexpr int kCompactThreads = 256;
constexpr int kNL = 12;

__global__ void __launch_bounds__(kCompactThreads)
k2_compact(Dims2 D, const uint8_t* m0, const uint8_t* m1, const uint8_t* m2, const uint8_t* l0, const uint8_t* l1,
           const uint8_t* l2, uint32_t* lists, Scal2* sc) {
    const int skip = sc->skip;  // tested after the mask and level loads
    __shared__ int wtot[kNL][kCompactThreads / 64];
    __shared__ int base[kNL];
    const long long N = (long long)D.Nx * D.Ny;
    const uint8_t* mk[3] = {m0, m1, m2};
    const uint8_t* lk[3] = {l0, l1, l2};
    int cnt[kNL];
#pragma unroll
    for (int q = 0; q < kNL; q++) cnt[q] = 0;
    uint32_t fl[16];
    Seg s;
    const bool ok = seg_of(D, (long long)blockIdx.x * blockDim.x + threadIdx.x, s) && s.j >= 0 && s.j <= D.Ny - 1;
    if (ok) {
        // every byte the segment needs, as independent loads: masks and levels
        // of the row and of the row below (16-byte aligned), and of the point
        // left of the segment (row j >= 0 has ghost rows and a ghost column
        // before it, so p0 - 1 and p0 - P lie inside the buffer)
        V16 mr[3], lr[3], mb[3], lb[3];
        uint8_t ml[3], ll[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            mr[k] = ld16(mk[k] + s.p0);
            lr[k] = ld16(lk[k] + s.p0);
            mb[k] = ld16(mk[k] + s.p0 - D.P);
            lb[k] = ld16(lk[k] + s.p0 - D.P);
            ml[k] = mk[k][s.p0 - 1];
            ll[k] = lk[k][s.p0 - 1];
        }
        const uint8_t top = uint8_t(D.maxlev);
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int i = s.i0 + e;
            uint32_t f = 0;
            if (i >= 0 && i <= D.Nx - 1) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    if (!mr[k].b[e]) continue;
                    if (lr[k].b[e] != top) { f |= 1u << (4 * k + 1); cnt[4 * k + 1]++; continue; }
                    f |= 1u << (4 * k);
                    cnt[4 * k]++;
                    const bool lm = e ? mr[k].b[e - 1] && lr[k].b[e - 1] == top : ml[k] && ll[k] == top;
                    const bool lw = i >= 1 && lm;
                    const bool bw = s.j >= 1 && mb[k].b[e] && lb[k].b[e] == top;
                    if (!lw) { f |= 1u << (4 * k + 2); cnt[4 * k + 2]++; }
                    if (!bw) { f |= 1u << (4 * k + 3); cnt[4 * k + 3]++; }
                }
            }
            fl[e] = f;
        }
    }
    if (skip) return;  // uniform over the block, before its barriers
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl[kNL];
#pragma unroll
    for (int q = 0; q < kNL; q++) {
        int x = cnt[q];
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        incl[q] = x;
        if (lane == 63) wtot[q][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < kNL) {
        int acc = 0;
        for (int w = 0; w < kCompactThreads / 64; w++) { const int x = wtot[threadIdx.x][w]; wtot[threadIdx.x][w] = acc; acc += x; }
        const int k = threadIdx.x >> 2, ty = threadIdx.x & 3;
        int* ctr = ty == 0 ? &sc->cntW[k] : ty == 1 ? &sc->cntC[k] : ty == 2 ? &sc->cntXS[k] : &sc->cntYS[k];
        base[threadIdx.x] = acc ? atomicAdd(ctr, acc) : 0;
    }
    __syncthreads();
    if (!ok) return;
    int pos[kNL];
#pragma unroll
    for (int q = 0; q < kNL; q++) pos[q] = base[q] + wtot[q][wid] + incl[q] - cnt[q];
    for (int e = 0; e < 16; e++) {
        const uint32_t f = fl[e];
        if (!f) continue;
        const uint32_t packed = (uint32_t(s.j) << 16) | uint32_t(s.i0 + e);
#pragma unroll
        for (int q = 0; q < kNL; q++)
            if (f & (1u << q)) lists[(long long)q * N + pos[q]++] = packed;
    }
}

// ---------------------------------------------------- CFL and LF alphas ----
// alpha_x(j) = max_i SRx, alpha_y(i) = max_j SRy over `mask` (exact maxima), and
// for stage 1 also max(SRx + SRy) for the CFL step (main_2d_euler(dmach).cpp:61-104).
// k2_alpha: a block takes a tile of 64 columns x kAlphaRows rows and writes its
// row, column and CFL maxima as partials.  The ghost-fill kernel that follows
// on the stream (k2_bc) reduces them into alpha_x, alpha_y and the CFL maximum.
// Maxima combine as unsigned bit patterns of non-negative doubles starting
// from 0, with partials that are not > 0 (NaN kept) replaced by 0: exact, and
// in every case the value that atomicMax of each positive partial into a
// zeroed array gives.
constexpr int kAlphaRows = 32;  // rows per block; each thread takes kAlphaRows/4 of them

__device__ __forceinline__ double bits_max(double a, double b) {
    return (unsigned long long)__double_as_longlong(b) > (unsigned long long)__double_as_longlong(a) ? b : a;
}
__device__ __forceinline__ double pos_or_zero(double v) { return (v <= 0) ? 0.0 : v; }

// bit-pattern maximum of p[0], p[stride], ..., p[(n-1) stride] and 0, with the
// loads issued eight at a time
__device__ __forceinline__ double strided_bits_max(const double* p, long long stride, int n) {
    double m = 0.0;
    int b = 0;
    for (; b + 8 <= n; b += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = p[(b + u) * stride];
#pragma unroll
        for (int u = 0; u < 8; u++) m = bits_max(m, v[u]);
    }
    for (; b < n; b++) m = bits_max(m, p[b * stride]);
    return m;
}

struct AlphaPart {
    double* px;  // [ntx][Ny]  row maxima of each tile column
    double* py;  // [nty][Nx]  column maxima of each tile row
    double* pv;  // [nty][ntx] CFL maxima per tile
    int ntx, nty;
};

__global__ void __launch_bounds__(256) k2_alpha(Dims2 D, const uint8_t* __restrict__ mask, const double* __restrict__ Q,
                                                AlphaPart P, double gamma, const Scal2* sc) {
    const int skip = sc->skip;  // tested after the mask loads
    constexpr int R = kAlphaRows / 4;
    __shared__ double colr[4][64];
    __shared__ double vr[4];
    const int tx = blockIdx.x, ty = blockIdx.y;
    const int i = tx * 64 + threadIdx.x;
    const int j0 = ty * kAlphaRows + threadIdx.y;
    // all mask loads, then all state loads, then the arithmetic
    uint8_t act[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = j0 + 4 * r;
        act[r] = (i < D.Nx && j < D.Ny) ? mask[gidx(D, i, j)] : 0;
    }
    if (skip) return;  // uniform over the block, before its barrier
    double q[R][4];
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (act[r]) {
            const long long p = gidx(D, i, j0 + 4 * r);
            q[r][0] = Q[p]; q[r][1] = Q[p + D.plane]; q[r][2] = Q[p + 2 * D.plane]; q[r][3] = Q[p + 3 * D.plane];
        } else {
            q[r][0] = 1; q[r][1] = 0; q[r][2] = 0; q[r][3] = 1;
        }
    }
    double cm = 0, vm = 0, sxr[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        double sx = 0;
        if (act[r]) {
            sx = spec_rad2x(gamma, q[r]);
            const double sy = spec_rad2y(gamma, q[r]);
            cm = t_max(cm, sy);
            vm = t_max(vm, sx + sy);
        }
        sxr[r] = sx;
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = j0 + 4 * r;
        const double rm = wave_max(sxr[r]);
        if (threadIdx.x == 0 && j < D.Ny) P.px[(long long)tx * D.Ny + j] = pos_or_zero(rm);
    }
    colr[threadIdx.y][threadIdx.x] = cm;
    const double wv = wave_max(vm);
    if (threadIdx.x == 0) vr[threadIdx.y] = wv;
    __syncthreads();
    if (threadIdx.y == 0) {
        const double c = t_max(t_max(colr[0][threadIdx.x], colr[1][threadIdx.x]), t_max(colr[2][threadIdx.x], colr[3][threadIdx.x]));
        if (i < D.Nx) P.py[(long long)ty * D.Nx + i] = pos_or_zero(c);
        if (threadIdx.x == 0) P.pv[ty * P.ntx + tx] = pos_or_zero(t_max(t_max(vr[0], vr[1]), t_max(vr[2], vr[3])));
    }
}

// ------------------------------------------------------------ ghost fill ----
__device__ __forceinline__ void put4(double* Q, long long plane, long long p, const double* v) {
    Q[p] = v[0]; Q[p + plane] = v[1]; Q[p + 2 * plane] = v[2]; Q[p + 3 * plane] = v[3];
}
__device__ __forceinline__ void get4(const double* Q, long long plane, long long p, double* v) {
    v[0] = Q[p]; v[1] = Q[p + plane]; v[2] = Q[p + 2 * plane]; v[3] = Q[p + 3 * plane];
}

// Ghost state on side 0..3 (left, right, bottom, top) in column i, from the
// state `src` of the boundary point it faces: a copy (transmissive,
// PROBLEM_QUADRANTS.h:123-203) or, for the double Mach reflection
// (BOUNDARY_MANAGER_DMACH::SetBoundaryCondition, PROBLEM_DOUBLEMACH.h:106-245,
// EOS at gamma = 1.4), the post-shock state on the left, a copy on the right,
// the reflecting wall (y-momentum negated) or the post-shock state below, and
// the post- or pre-shock state above as the shock has moved by time T.
__device__ __forceinline__ void ghost_state(int kind, int side, int i, double x0, double dx, double T,
                                            const double* src, double* v) {
    if (kind == AWCM_BC_TRANSMISSIVE) {
        for (int k = 0; k < 4; k++) v[k] = src[k];
        return;
    }
    const double sqrt3 = sqrt(3.0);
    if (side == 0) {
        const double rho = 8, u = 8.25 * sqrt(3.0) * 0.5, vv = -8.25 * 0.5, pr = 116.5;
        v[0] = rho; v[1] = rho * u; v[2] = rho * vv; v[3] = eos_E2(1.4, rho, u, vv, pr);
    } else if (side == 1) {
        for (int k = 0; k < 4; k++) v[k] = src[k];
    } else if (side == 2) {
        const double x = x0 + (i - 0) * dx;
        if (x > 1.0 / 6.0) {
            for (int k = 0; k < 4; k++) v[k] = src[k];
            v[2] = -v[2];
        } else {
            v[0] = 8; v[1] = 8 * 8.25 * sqrt3 * 0.5; v[2] = -8 * 8.25 * 0.5;
            v[3] = eos_E2(1.4, 8, 8.25 * sqrt3 * 0.5, -8.25 * 0.5, 116.5);
        }
    } else {
        const double x = x0 + (i - 0) * dx;
        if (x < 1.0 / 6.0 + Div(sqrt3)(1 + 20 * T)) {
            v[0] = 8; v[1] = 8 * 8.25 * sqrt3 * 0.5; v[2] = -8 * 8.25 * 0.5;
            v[3] = eos_E2(1.4, 8, 8.25 * sqrt3 * 0.5, -8.25 * 0.5, 116.5);
        } else {
            v[0] = 1.4; v[1] = 1.4 * 0; v[2] = 1.4 * 0; v[3] = eos_E2(1.4, 1.4, 0, 0, 1);
        }
    }
}

// Ghost fill of one field at time t; with compute_dt, block 0 first turns the
// CFL maximum into dt with the output-time clip (main_2d_euler(dmach).cpp:76-83).
// Blocks nbc.. of the grid reduce the alpha partials k2_alpha left (one row or
// column per thread); with compute_dt, block 0 first reduces the CFL partials
// and turns the maximum into dt with the output-time clip
// (main_2d_euler(dmach).cpp:76-83).  The other blocks fill the ghosts.
__global__ void __launch_bounds__(256) k2_bc(Dims2 D, double* Q, int kind, double x0, double dx, int compute_dt,
                                             const double* pt, AlphaPart P, double* ax, double* ay, int nbc,
                                             double* zero_next, Scal2* sc) {
    const int skip = sc->skip;  // tested after the loads, before any store
    if ((int)blockIdx.x >= nbc) {
        const int r = (blockIdx.x - nbc) * blockDim.x + threadIdx.x;
        // the next two stages' alpha arrays, which their predecessors' stage
        // updates fill with atomic maxima
        if (zero_next && !skip && r < D.Nx + D.Ny) {
            zero_next[r] = 0.0;
            zero_next[D.Nx + D.Ny + r] = 0.0;
        }
        if (r < D.Ny) {
            const double m = strided_bits_max(P.px + r, D.Ny, P.ntx);
            if (!skip) ax[r] = m;
        } else if (r < D.Ny + D.Nx) {
            const double m = strided_bits_max(P.py + (r - D.Ny), D.Nx, P.nty);
            if (!skip) ay[r - D.Ny] = m;
        }
        return;
    }
    if (compute_dt && blockIdx.x == 0) {
        __shared__ double red[256];
        double m = 0.0;
        for (int k = threadIdx.x; k < P.ntx * P.nty; k += blockDim.x) m = bits_max(m, P.pv[k]);
        red[threadIdx.x] = m;
        __syncthreads();
        if (threadIdx.x == 0 && !skip) {
            double vmax = 0.0;
            for (int k = 0; k < (int)blockDim.x; k++) vmax = bits_max(vmax, red[k]);
            sc->vmax = (unsigned long long)__double_as_longlong(vmax);
            double dt = Div(vmax)(sc->cfl * dx);
            int dp = 0;
            if (sc->cnt < sc->n_print && sc->t + dt > pt[sc->cnt]) { dt = pt[sc->cnt] - sc->t; dp = 1; }
            sc->dt = dt;
            sc->do_print = dp;
            if (!(vmax > 0) || !(vmax < INFINITY)) sc->flag |= 2;
        }
    }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nl = G * D.Ny, nb = G * D.Nx;
    const double T = sc->t;
    int side, i, j;
    if (t < nl) { side = 0; j = t / G; i = -1 - t % G; }
    else if (t < 2 * nl) { side = 1; j = (t - nl) / G; i = D.Nx + (t - nl) % G; }
    else if (t < 2 * nl + nb) { side = 2; i = (t - 2 * nl) / G; j = -1 - (t - 2 * nl) % G; }
    else if (t < 2 * nl + 2 * nb) { side = 3; i = (t - 2 * nl - nb) / G; j = D.Ny + (t - 2 * nl - nb) % G; }
    else return;
    // the boundary point the ghost is filled from
    const int si = side == 0 ? 0 : side == 1 ? D.Nx - 1 : i;
    const int sj = side == 2 ? 0 : side == 3 ? D.Ny - 1 : j;
    double src[4], v[4];
    get4(Q, D.plane, gidx(D, si, sj), src);
    ghost_state(kind, side, i, x0, dx, T, src, v);
    if (!skip) put4(Q, D.plane, gidx(D, i, j), v);
}

// ------------------------------------------------------------ RK stages ----
// S = 1: Q2 = Q - dt/h r(Q);  S = 2: Q1 = 0.75 Q + 0.25 (Q2 - dt/h r(Q2));
// S = 3: Q0 = Q1 - dt/h r(Q1)  (main_2d_euler(dmach).cpp:162, 242, 321),
// folded at once into the RK combine Q = (Q + 2 Q0) / 3 (:325-332): Qo is Q.
template <int S>
__device__ __forceinline__ void stage_out(const Dims2& D, long long p, const double* r, double h, double dt,
                                          const double* Qa, const double* Qs, double* Qo, double* out) {
    const double dth = Div(h)(dt);
    const Div third(3.0);
    for (int k = 0; k < 4; k++) {
        const long long o = p + k * D.plane;
        if (S == 1) out[k] = Qs[o] - dth * r[k];
        else if (S == 2) out[k] = 0.75 * Qa[o] + 0.25 * (Qs[o] - dth * r[k]);
        else out[k] = Qs[o] - dth * r[k];
        if (S == 3) Qo[o] = third(Qo[o] + 2.0 * out[k]);
        else Qo[o] = out[k];
    }
}

// The ghosts a boundary point (i, j) of the new field Qo feeds (k2_bc's fill,
// done by the thread that writes the point): every boundary point is in every
// stage mask (the band of k2_band_snap).
__device__ __forceinline__ void fill_ghosts_from(const Dims2& D, int kind, double x0, double dx, double T, int i, int j,
                                                 const double* q, double* Qo) {
    double v[4];
    if (i == 0) {
        ghost_state(kind, 0, -1, x0, dx, T, q, v);
        for (int c = 0; c < G; c++) put4(Qo, D.plane, gidx(D, -1 - c, j), v);
    }
    if (i == D.Nx - 1) {
        ghost_state(kind, 1, D.Nx, x0, dx, T, q, v);
        for (int c = 0; c < G; c++) put4(Qo, D.plane, gidx(D, D.Nx + c, j), v);
    }
    if (j == 0) {
        ghost_state(kind, 2, i, x0, dx, T, q, v);
        for (int c = 0; c < G; c++) put4(Qo, D.plane, gidx(D, i, -1 - c), v);
    }
    if (j == D.Ny - 1) {
        ghost_state(kind, 3, i, x0, dx, T, q, v);
        for (int c = 0; c < G; c++) put4(Qo, D.plane, gidx(D, i, D.Ny + c), v);
    }
}

// WENO-Z/Roe face fluxes of one stage, each face once.  Face slot p = gidx(i,j)
// holds the face to the right of (FX) / above (FY) point (i, j).  Faces needed:
// the right and top faces of every finest-level point, plus the left face of
// each x-run start and the bottom face of each y-run start.  One face per
// thread; the flux is WENOZ_EULER_SOLVER_2D::ComputeFluxX/Y (WENOZ.h:236-320) of
// the six states around the face -- the same call the reference makes, once
// per face instead of once per adjacent point.
__global__ void __launch_bounds__(128) k2_stage_faces(Dims2 D, const uint32_t* __restrict__ lw,
                                                      const uint32_t* __restrict__ lxs, const uint32_t* __restrict__ lys,
                                                      int which, const double* __restrict__ Qs, double* __restrict__ FX,
                                                      double* __restrict__ FY, const double* __restrict__ ax,
                                                      const double* __restrict__ ay, Scal2* sc) {
    if (sc->skip) return;
    const int nW = sc->cntW[which], nX = sc->cntXS[which], nY = sc->cntYS[which];
    const int nx = nW + nX, ntot = nx + nW + nY;
    // plain grid stride: a contiguous share of the list per XCD cuts the L2-miss
    // traffic by about a tenth but unbalances x- and y-faces, and this kernel is
    // bound by FP64 issue, not by memory
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntot; t += gridDim.x * blockDim.x) {
        const bool xdir = t < nx;
        const int u = xdir ? t : t - nx;
        uint32_t e;
        int shift;  // 0: the face after the point; -1: the face before it
        if (xdir) { if (u < nW) { e = lw[u]; shift = 0; } else { e = lxs[u - nW]; shift = -1; } }
        else { if (u < nW) { e = lw[u]; shift = 0; } else { e = lys[u - nW]; shift = -1; } }
        const int i = int(e & 0xffffu), j = int(e >> 16);
        double s[6][4], f[4];
        if (xdir) {
            const long long p = gidx(D, i + shift, j);  // face between p and p + 1
            for (int q = 0; q < 6; q++) get4(Qs, D.plane, p + (q - 2), s[q]);
            face_flux2<0>(kFluxGamma, s, ax[j], f);
            put4(FX, D.plane, p, f);
        } else {
            const long long p = gidx(D, i, j + shift);  // face between p and p + P
            for (int q = 0; q < 6; q++) get4(Qs, D.plane, p + (long long)(q - 2) * D.P, s[q]);
            face_flux2<1>(kFluxGamma, s, ay[i], f);
            put4(FY, D.plane, p, f);
        }
    }
}

// Stage update of every point of the stage mask: finest-level points difference
// their face fluxes (dF = right - left, dG = top - bottom: main_2d_euler(dmach).cpp
// :123-128); coarser points take central differences of the physical flux at
// spacing d with the one-sided variants of AWCM_2D.h:166-187.
//
// With gkind >= 0 (stages 1 and 2) the kernel also fills the ghosts of its
// output, the next stage's input (k2_bc's fill, from the boundary points it
// writes).  Stage 3 folds the RK combine into its output (stage_out<3>).
//
// With `nax` (stages 1 and 2), the kernel also forms the next stage's LF
// alphas: the points it writes are exactly the next stage's mask (mask2 for
// stage 1, mask1 for stage 2) and their new states are the next stage's input,
// so alpha_x(j) = max SRx over row j and alpha_y(i) = max SRy over column i
// (main_2d_euler(dmach).cpp:167-183, 247-263) are gathered here as bit-pattern
// maxima of non-negative doubles (exact and order-free) into arrays an earlier
// kernel of the step zeroed: one atomic per row present in a wavefront, one per
// point for the columns.  A wavefront's lanes run the loop in lockstep so the
// row reduction sees every lane.
__device__ __forceinline__ void next_alpha(bool valid, int i, int j, const double* q, double gamma,
                                           double* __restrict__ nax, double* __restrict__ nay) {
    double sx = 0, sy = 0;
    if (valid) {
        sx = pos_or_zero(spec_rad2x(gamma, q));
        sy = pos_or_zero(spec_rad2y(gamma, q));
    }
    const int lane = threadIdx.x & 63;
    unsigned long long live = __ballot(valid);
    while (live) {
        const int leader = __builtin_ctzll(live);
        const int jl = __shfl(j, leader, 64);
        const bool in = valid && j == jl;
        const unsigned long long grp = __ballot(in);
        const double m = wave_max(in ? sx : 0.0);
        if (lane == leader) atomic_max_pos(reinterpret_cast<unsigned long long*>(nax + jl), m);
        live &= ~grp;
    }
    if (valid) atomic_max_pos(reinterpret_cast<unsigned long long*>(nay + i), sy);
}

template <int S>
__global__ void __launch_bounds__(256) k2_stage_update(Dims2 D, const uint32_t* __restrict__ lw,
                                                       const uint32_t* __restrict__ lc, int which,
                                                       const uint8_t* __restrict__ lvl, const double* __restrict__ Qa,
                                                       const double* __restrict__ Qs, double* __restrict__ Qo,
                                                       const double* __restrict__ FX, const double* __restrict__ FY,
                                                       double gamma, double gdx, double* __restrict__ nax,
                                                       double* __restrict__ nay, int gkind, double x0, Scal2* sc) {
    if (sc->skip) return;
    const int nW = sc->cntW[which], nC = sc->cntC[which];
    const double dt = sc->dt, T = sc->t;
    const int stride = gridDim.x * blockDim.x;
    // wave-uniform trip count: every lane of a wavefront runs every iteration
    for (int t0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63); t0 < nW + nC; t0 += stride) {
        const int t = t0 + (threadIdx.x & 63);
        bool valid = t < nW + nC;
        int i = 0, j = 0;
        double qn[4];
        if (valid && t < nW) {
            const uint32_t e = lw[t];
            i = int(e & 0xffffu), j = int(e >> 16);
            const long long p = gidx(D, i, j);
            double a[4], b[4], c[4], d[4], r[4];
            get4(FX, D.plane, p - 1, a);
            get4(FX, D.plane, p, b);
            get4(FY, D.plane, p - D.P, c);
            get4(FY, D.plane, p, d);
            for (int k = 0; k < 4; k++) r[k] = (b[k] - a[k]) + (d[k] - c[k]);
            stage_out<S>(D, p, r, gdx * 1, dt, Qa, Qs, Qo, qn);
        } else if (valid) {
            const uint32_t e = lc[t - nW];
            i = int(e & 0xffffu), j = int(e >> 16);
            const long long p = gidx(D, i, j);
            const int lev = lvl[p];
            const int d = 1 << (D.L - 1 - lev);
            // a closed mask keeps every used stencil point inside the ghost band
            const bool okx = (i <= d) ? (i - d >= -G && i + 2 * d <= D.Nx - 1 + G)
                                      : (i - 2 * d >= -G && i + ((i >= D.Nx - 1 - d) ? d : 2 * d) <= D.Nx - 1 + G);
            const bool oky = (j <= d) ? (j - d >= -G && j + 2 * d <= D.Ny - 1 + G)
                                      : (j - 2 * d >= -G && j + ((j >= D.Ny - 1 - d) ? d : 2 * d) <= D.Ny - 1 + G);
            if (!okx || !oky) {
                sc->flag |= 1;
                valid = false;
            } else {
                double q[4], fll[4], fl[4], fr[4], frr[4], dF[4], dG[4];
                const long long sx = d, sy = (long long)d * D.P;
                if (i <= d) {
                    get4(Qs, D.plane, p - sx, q); flux_f2<0>(gamma, q, fl);
                    get4(Qs, D.plane, p + sx, q); flux_f2<0>(gamma, q, fr);
                    get4(Qs, D.plane, p + 2 * sx, q); flux_f2<0>(gamma, q, frr);
                    for (int k = 0; k < 4; k++) dF[k] = kD1 * (fr[k] - fl[k]) + kD2 * (frr[k] - fl[k]);
                } else if (i >= D.Nx - 1 - d) {
                    get4(Qs, D.plane, p - 2 * sx, q); flux_f2<0>(gamma, q, fll);
                    get4(Qs, D.plane, p - sx, q); flux_f2<0>(gamma, q, fl);
                    get4(Qs, D.plane, p + sx, q); flux_f2<0>(gamma, q, fr);
                    for (int k = 0; k < 4; k++) dF[k] = kD1 * (fr[k] - fl[k]) + kD2 * (fr[k] - fll[k]);
                } else {
                    get4(Qs, D.plane, p - 2 * sx, q); flux_f2<0>(gamma, q, fll);
                    get4(Qs, D.plane, p - sx, q); flux_f2<0>(gamma, q, fl);
                    get4(Qs, D.plane, p + sx, q); flux_f2<0>(gamma, q, fr);
                    get4(Qs, D.plane, p + 2 * sx, q); flux_f2<0>(gamma, q, frr);
                    for (int k = 0; k < 4; k++) dF[k] = kD1 * (fr[k] - fl[k]) + kD2 * (frr[k] - fll[k]);
                }
                if (j <= d) {
                    get4(Qs, D.plane, p - sy, q); flux_f2<1>(gamma, q, fl);
                    get4(Qs, D.plane, p + sy, q); flux_f2<1>(gamma, q, fr);
                    get4(Qs, D.plane, p + 2 * sy, q); flux_f2<1>(gamma, q, frr);
                    for (int k = 0; k < 4; k++) dG[k] = kD1 * (fr[k] - fl[k]) + kD2 * (frr[k] - fl[k]);
                } else if (j >= D.Ny - 1 - d) {
                    get4(Qs, D.plane, p - 2 * sy, q); flux_f2<1>(gamma, q, fll);
                    get4(Qs, D.plane, p - sy, q); flux_f2<1>(gamma, q, fl);
                    get4(Qs, D.plane, p + sy, q); flux_f2<1>(gamma, q, fr);
                    for (int k = 0; k < 4; k++) dG[k] = kD1 * (fr[k] - fl[k]) + kD2 * (fr[k] - fll[k]);
                } else {
                    get4(Qs, D.plane, p - 2 * sy, q); flux_f2<1>(gamma, q, fll);
                    get4(Qs, D.plane, p - sy, q); flux_f2<1>(gamma, q, fl);
                    get4(Qs, D.plane, p + sy, q); flux_f2<1>(gamma, q, fr);
                    get4(Qs, D.plane, p + 2 * sy, q); flux_f2<1>(gamma, q, frr);
                    for (int k = 0; k < 4; k++) dG[k] = kD1 * (fr[k] - fl[k]) + kD2 * (frr[k] - fll[k]);
                }
                double r[4];
                for (int k = 0; k < 4; k++) r[k] = dF[k] + dG[k];
                stage_out<S>(D, p, r, gdx * d, dt, Qa, Qs, Qo, qn);
            }
        }
        if (gkind >= 0 && valid && (i == 0 || i == D.Nx - 1 || j == 0 || j == D.Ny - 1))
            fill_ghosts_from(D, gkind, x0, gdx, T, i, j, qn, Qo);
        if (nax) next_alpha(valid, i, j, qn, gamma, nax, nay);
    }
}

// RemoveResidueEuler (AWCM_MANAGER_2D.h:267-280) on the interior; with `advance`
// (the end of a step, whose RK combine stage 3 already did) thread 0 also
// advances t and the step counters.  Between steps Q is zero at every interior
// point outside mask0, and within a step only points of mask3 (inverse
// transform, interpolation) or of the previous mask0 (mask_temp) can become
// non-zero, so the pass only writes there -- unless the host replaced Q or
// mask0 (residue_full).
__global__ void k2_residue(Dims2 D, const uint8_t* __restrict__ m0, const uint8_t* __restrict__ m3,
                           const uint8_t* __restrict__ mtmp, double* __restrict__ Q, int advance, Scal2* sc) {
    const int skip = sc->skip;  // tested after the mask load
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const long long p = gidx(D, i < D.Nx ? i : D.Nx - 1, j);
    const uint8_t a0 = m0[p];
    if (skip) return;
    if (advance && i == 0 && j == 0) {
        sc->t = sc->t + sc->dt;
        if (sc->do_print) sc->cnt = sc->cnt + 1;
        sc->steps += 1;
        sc->sum_active0 += (long long)sc->cntW[0] + sc->cntC[0];
        sc->sumW += (long long)sc->cntW[0] + sc->cntW[1] + sc->cntW[2];
        sc->sumC += (long long)sc->cntC[0] + sc->cntC[1] + sc->cntC[2];
        for (int k = 0; k < 3; k++) sc->sumF += 2LL * sc->cntW[k] + sc->cntXS[k] + sc->cntYS[k];
    }
    if (i >= D.Nx || a0) return;
    if (i >= 1 && i <= D.Nx - 2 && j >= 1 && j <= D.Ny - 2 && (sc->residue_full || m3[p] || mtmp[p])) {
        for (int q = 0; q < 4; q++) Q[p + q * D.plane] = 0.0;
    }
}

// -------------------------------------------------- scalar box advection ----
// The scalar driver's iteration after its Prepare (Wavelet/main_2d_advect.cpp:
// 50-198) on plane 0 of Q, velocity (1, 1), with the driver's three stage
// fields A2, A1, A0 (one plane each, kept between iterations as the driver
// keeps Q_2, Q_1, Q_0).

// BOUNDARY_MANAGER::SetBoundaryCondition of PROBLEM_BOX_ADVECTION.h:78-151:
// periodic x ghosts of rows 0..Ny-1 and y ghosts of columns 0..Nx-1 (corners
// untouched).  Every source is an interior point, so no ghost reads another.
__global__ void k2_periodic(Dims2 D, double* __restrict__ F) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nxs = G * D.Ny, nys = G * D.Nx;
    if (t < 2 * nxs) {
        const int r = t % nxs, j = r / G, c = r % G;
        if (t < nxs) F[gidx(D, -1 - c, j)] = F[gidx(D, D.Nx - 1 - c, j)];
        else F[gidx(D, D.Nx + c, j)] = F[gidx(D, c, j)];
    } else if (t < 2 * nxs + 2 * nys) {
        const int u = t - 2 * nxs, r = u % nys, i = r / G, c = r % G;
        if (u < nys) F[gidx(D, i, -1 - c)] = F[gidx(D, i, D.Ny - 1 - c)];
        else F[gidx(D, i, D.Ny + c)] = F[gidx(D, i, c)];
    }
}

// One stage over the mask's points (finest points dense per row, one thread
// each): dF + dG from WENO5 right reconstructions on the finest level
// (main_2d_advect.cpp:63-73) or AWCM_2D::Differentiate and its one-sided forms
// at spacing d below it (:74-103, AWCM_2D.h:95-110), then
//   S = 1: out = F - dt/h (dF + dG)                          (:104, mask2)
//   S = 2: out = 0.75 Qa + 0.25 (F - dt/h (dF + dG))         (:149, mask1)
//   S = 3: out = F - dt/h (dF + dG)                          (:194, mask0)
// with h = dx * d.  A coarse stencil that would leave the ghosted grid (the
// reference reads outside its arrays there) sets the range flag.
template <int S>
__global__ void __launch_bounds__(256) k2_advect_stage(Dims2 D, const uint8_t* __restrict__ mask,
                                                       const uint8_t* __restrict__ lvl, const double* __restrict__ Qa,
                                                       const double* __restrict__ F, double* __restrict__ out, double dt,
                                                       double gdx, Scal2* sc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= D.Nx) return;
    const long long p = gidx(D, i, j);
    if (!mask[p]) return;
    const int lev = lvl[p];
    const int d = 1 << (D.L - 1 - lev);
    double r;
    if (lev == D.maxlev) {
        const long long P = D.P;
        double a = wenojs_right(F[p - 3], F[p - 2], F[p - 1], F[p], F[p + 1]);
        double b = wenojs_right(F[p - 2], F[p - 1], F[p], F[p + 1], F[p + 2]);
        const double dF = b - a;
        a = wenojs_right(F[p - 3 * P], F[p - 2 * P], F[p - P], F[p], F[p + P]);
        b = wenojs_right(F[p - 2 * P], F[p - P], F[p], F[p + P], F[p + 2 * P]);
        const double dG = b - a;
        r = dF + dG;
    } else {
        const bool lx = i <= d, rx = !lx && i >= D.Nx - 1 - d;
        const bool ly = j <= d, ry = !ly && j >= D.Ny - 1 - d;
        const bool okx = i - (lx ? d : 2 * d) >= -G && i + (rx ? d : 2 * d) <= D.Nx - 1 + G;
        const bool oky = j - (ly ? d : 2 * d) >= -G && j + (ry ? d : 2 * d) <= D.Ny - 1 + G;
        if (!okx || !oky) { sc->flag |= 1; return; }
        const long long sx = d, sy = (long long)d * D.P;
        double dF, dG;
        if (lx) dF = kD1 * (F[p + sx] - F[p - sx]) + kD2 * (F[p + 2 * sx] - F[p - sx]);
        else if (rx) dF = kD1 * (F[p + sx] - F[p - sx]) + kD2 * (F[p + sx] - F[p - 2 * sx]);
        else dF = kD1 * (F[p + sx] - F[p - sx]) + kD2 * (F[p + 2 * sx] - F[p - 2 * sx]);
        if (ly) dG = kD1 * (F[p + sy] - F[p - sy]) + kD2 * (F[p + 2 * sy] - F[p - sy]);
        else if (ry) dG = kD1 * (F[p + sy] - F[p - sy]) + kD2 * (F[p + sy] - F[p - 2 * sy]);
        else dG = kD1 * (F[p + sy] - F[p - sy]) + kD2 * (F[p + 2 * sy] - F[p - 2 * sy]);
        r = dF + dG;
    }
    const double h = gdx * d;
    if (S == 2) out[p] = 0.75 * Qa[p] + 0.25 * (F[p] - Div(h)(dt) * r);
    else out[p] = F[p] - Div(h)(dt) * r;
}

// The combine on mask0 (main_2d_advect.cpp:190-197) and PostProcess:
// AWCM_MANAGER_2D::RemoveResidue on the interior (AWCM_MANAGER_2D.h:257-266).
__global__ void k2_advect_combine(Dims2 D, const uint8_t* __restrict__ m0, double* __restrict__ Q,
                                  const double* __restrict__ A0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= D.Nx) return;
    const long long p = gidx(D, i, j);
    if (m0[p]) Q[p] = (1.0 / 3.0 * Q[p] + 2.0 / 3.0 * A0[p]);
    else if (i >= 1 && i <= D.Nx - 2 && j >= 1 && j <= D.Ny - 2) Q[p] = 0.0;
}

// AoS, reference pitch (Nx+6) <-> SoA planes, device pitch P
__global__ void k2_aos_to_soa(const double* __restrict__ aos, double* __restrict__ soa, long long n, int PR, int P,
                              long long plane) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long long d = (t % PR) + (long long)P * (t / PR);
    for (int k = 0; k < 4; k++) soa[d + k * plane] = aos[t * 4 + k];
}
__global__ void k2_soa_to_aos(const double* __restrict__ soa, double* __restrict__ aos, long long n, int PR, int P,
                              long long plane) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long long d = (t % PR) + (long long)P * (t / PR);
    for (int k = 0; k < 4; k++) aos[t * 4 + k] = soa[d + k * plane];
}
// level field (ghosted device layout) -> int32, index i + Nx*j
__global__ void k2_widen_levels(Dims2 D, const uint8_t* __restrict__ src, int32_t* __restrict__ dst) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)D.Nx * D.Ny) return;
    const int i = int(t % D.Nx), j = int(t / D.Nx);
    dst[t] = src[gidx(D, i, j)];
}

// ---------------------------------- AWCM_MANAGER_2D statics on one mask ----
// AWCM_MANAGER_2D::ComputeLevel (AWCM_MANAGER_2D.h:215-241) at every finest
// point: 0 where the mask is off; else the finest level with an active
// 4-neighbour at that level's spacing; unchanged when there is none.  A
// neighbour outside the ghosted domain sets the error flag.
__global__ void k2_compute_level(Dims2 D, const uint8_t* __restrict__ mask, uint8_t* __restrict__ lvl, Scal2* sc) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)D.Nx * D.Ny) return;
    const int i = int(t % D.Nx), j = int(t / D.Nx);
    const long long p = gidx(D, i, j);
    if (!mask[p]) { lvl[p] = 0; return; }
    int d = 1;
    for (int l = D.maxlev; l >= 0; l--) {
        if (i - d < -G || i + d > D.Nx - 1 + G || j - d < -G || j + d > D.Ny - 1 + G) { sc->flag |= 1; return; }
        if (mask[p - d] || mask[p + d] || mask[p - (long long)d * D.P] || mask[p + (long long)d * D.P]) {
            lvl[p] = (uint8_t)l;
            return;
        }
        d *= 2;
    }
}

// AWCM_MANAGER_2D::AddPointForDifferentiation (AWCM_MANAGER_2D.h:160-213) at
// every interior point of the snapshot, with the caller's levels.
__global__ void k2_add_point(Dims2 D, const uint8_t* __restrict__ snap, uint8_t* __restrict__ mask,
                             const uint8_t* __restrict__ lvl, Scal2* sc) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ni = D.Nx - 2, nj = D.Ny - 2;
    if (t >= (long long)ni * nj) return;
    const int i = 1 + int(t % ni), j = 1 + int(t / ni);
    const long long p = gidx(D, i, j);
    if (!snap[p]) return;
    const int lev = lvl[p];
    const int dd = 1 << (D.L - 1 - lev);  // grid[lev].dx
    if (lev == D.maxlev) {
        for (int c = -3; c <= 3; c++) { mset(D, mask, i + c, j, sc); mset(D, mask, i, j + c, sc); }
        return;
    }
    int lo, hi;
    if (i - 2 * dd < 0) lo = -1, hi = 2;
    else if (i + 2 * dd > D.Nx - 1) lo = -2, hi = 1;
    else lo = -2, hi = 2;
    for (int c = lo; c <= hi; c++) mset(D, mask, i + dd * c, j, sc);
    if (j - 2 * dd < 0) lo = -1, hi = 2;
    else if (j + 2 * dd > D.Ny - 1) lo = -2, hi = 1;
    else lo = -2, hi = 2;
    for (int c = lo; c <= hi; c++) mset(D, mask, i, j + dd * c, sc);
}

// AWCM_MANAGER_2D::RemoveResidue / RemoveResidueEuler (AWCM_MANAGER_2D.h:
// 257-280) on an AoS field of ncomp components in the reference layout (row
// pitch PR): every component zeroed at the interior points outside the mask.
__global__ void k2_residue_aos(Dims2 D, double* __restrict__ q, int ncomp, const uint8_t* __restrict__ mask) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ni = D.Nx - 2, nj = D.Ny - 2;
    if (t >= (long long)ni * nj) return;
    const int i = 1 + int(t % ni), j = 1 + int(t / ni);
    if (mask[gidx(D, i, j)]) return;
    double* v = q + ((long long)(i + G) + (long long)D.PR * (j + G)) * ncomp;
    for (int c = 0; c < ncomp; c++) v[c] = 0.0;
}

inline unsigned nblk(long long n, int b) { return unsigned((n + b - 1) / b); }

// Host mirror of lattice(): thread counts of the lattice kernels.
Lat lat_h(const Dims2& D, int lev, int kind) {
    const int m = 1 << (D.L - 1 - lev), h = m >> 1;
    Lat g;
    g.dx = g.dy = m;
    if (kind == LAT_D1) { g.is = h; g.ie = D.Nx - 1 - h; g.js = m; g.je = D.Ny - 1 - m; }
    else if (kind == LAT_D2) { g.is = m; g.ie = D.Nx - 1 - m; g.js = h; g.je = D.Ny - 1 - h; }
    else if (kind == LAT_D3) { g.is = h; g.ie = D.Nx - 1 - h; g.js = h; g.je = D.Ny - 1 - h; }
    else { g.is = m; g.ie = D.Nx - 1 - m; g.js = m; g.je = D.Ny - 1 - m; }
    g.ni = (g.ie - g.is) / m + 1;
    g.nj = (g.je - g.js) / m + 1;
    return g;
}

Lat mklat(int is, int js, int ie, int je, int ires, int jres) {
    Lat g;
    g.is = is; g.js = js; g.ie = ie; g.je = je;
    g.dx = (ie - is) / (ires - 1);
    g.dy = (je - js) / (jres - 1);
    g.ni = (ie - is) / g.dx + 1;
    g.nj = (je - js) / g.dy + 1;
    return g;
}

}  // namespace

// ============================================================== host side ====

Solver2D::~Solver2D() {
    if (stream) hipStreamSynchronize(stream);
    if (gexec) hipGraphExecDestroy(gexec);
    if (graph) hipGraphDestroy(graph);
    for (hipEvent_t e : ev_pool) hipEventDestroy(e);
    for (void* p : allocs) hipFree(p);
    if (stream) hipStreamDestroy(stream);
    if (s_int) hipStreamDestroy(s_int);
    for (hipEvent_t e : {ev_m3, ev_int})
        if (e) hipEventDestroy(e);
}

int Solver2D::dalloc(void* pp, size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return AWCM_ENOMEM;
    allocs.push_back(p);
    *static_cast<void**>(pp) = p;
    return AWCM_OK;
}

int Solver2D::init(int L, int nxc_, int nyc_, double g, double dx_, double x0_, int bck) {
    if (L < 2 || L > 14 || nxc_ < 5 || nyc_ < 5) return AWCM_EINVAL;
    if (!(dx_ > 0) || !(g > 1)) return AWCM_EINVAL;
    if (bck != AWCM_BC_TRANSMISSIVE && bck != AWCM_BC_DOUBLE_MACH) return AWCM_EINVAL;
    if ((long long)(nxc_ - 1) * (1 << (L - 1)) + 1 > 65535 || (long long)(nyc_ - 1) * (1 << (L - 1)) + 1 > 65535)
        return AWCM_EINVAL;
    nxc = nxc_; nyc = nyc_; gamma = g; dx = dx_; x0 = x0_; bc_kind = bck;
    D.L = L;
    D.maxlev = L - 1;
    D.Nx = (nxc - 1) * (1 << (L - 1)) + 1;
    D.Ny = (nyc - 1) * (1 << (L - 1)) + 1;
    // Row pitch padded to 16 points so every mask row is 16-byte aligned and
    // every state row 128-byte aligned.  No stencil of a closed mask reaches past
    // the ghost band, so the reference's wrap-around pitch (Nx+6) need not be kept
    // on the device; the C ABI converts at upload/download.
    D.PR = D.Nx + 2 * G;
    D.P = (D.PR + 15) / 16 * 16;
    D.plane = (long long)D.P * (D.Ny + 2 * G);
    // MRA_GRID_INT_2D::Initialize (MRA_GRID_UNIFORM_2D.h:87-134)
    const int Nx = D.Nx, Ny = D.Ny;
    for (int lev = L - 1; lev >= 0; lev--) {
        const int m = 1 << (L - lev - 1);
        if (lev == L - 1) {
            grd[lev] = mklat(0, 0, Nx - 1, Ny - 1, (nxc - 1) * (1 << lev) + 1, (nyc - 1) * (1 << lev) + 1);
            continue;
        }
        grd[lev] = mklat(m, m, (Nx - 1) - m, (Ny - 1) - m, (nxc - 1) * (1 << lev) + 1 - 2, (nyc - 1) * (1 << lev) + 1 - 2);
        sub[lev][0] = mklat(int(0.5 * m), m, int(Nx - 0.5 * m - 1), Ny - m - 1, int(Nx / m), int(Ny / m - 1));
        sub[lev][1] = mklat(m, int(0.5 * m), Nx - m - 1, int(Ny - 0.5 * m - 1), int(Nx / m - 1), int(Ny / m));
        sub[lev][2] = mklat(int(0.5 * m), int(0.5 * m), int(Nx - 0.5 * m - 1), int(Ny - 0.5 * m - 1), int(Nx / m), int(Ny / m));
    }
    // the closed-form lattices used by the kernels equal the reference's ranges
    for (int lev = 0; lev < L - 1; lev++) {
        const Lat r[4] = {sub[lev][0], sub[lev][1], sub[lev][2], grd[lev]};
        for (int k = 0; k < 4; k++) {
            const Lat c = lat_h(D, lev, k);
            if (c.is != r[k].is || c.ie != r[k].ie || c.js != r[k].js || c.je != r[k].je || c.dx != r[k].dx ||
                c.dy != r[k].dy || c.ni != r[k].ni || c.nj != r[k].nj)
                return AWCM_EINVAL;
        }
    }
    set_tail(0);
    wavelet_weights(W.w);
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s_int, hipStreamNonBlocking) != hipSuccess)
        return AWCM_EDEVICE;
    for (hipEvent_t* e : {&ev_m3, &ev_int})
        if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) return AWCM_EDEVICE;
    const size_t pb = sizeof(double) * 4 * size_t(D.plane);
    const size_t mb = size_t(D.plane);
    const size_t nl = size_t(Nx) * Ny;
    int st;
    if ((st = dalloc(&Q, pb)) || (st = dalloc(&Q1, pb)) || (st = dalloc(&Q2, pb))) return st;
    if ((st = dalloc(&mt, mb)) || (st = dalloc(&snap, mb)) || (st = dalloc(&tmp, mb)) || (st = dalloc(&work, mb)))
        return st;
    for (int k = 0; k < 4; k++) if ((st = dalloc(&m[k], mb)) || (st = dalloc(&lv[k], mb))) return st;
    if ((st = dalloc(&alpha, sizeof(double) * 3 * (Nx + Ny)))) return st;
    {
        const size_t ntx = (size_t(Nx) + 63) / 64, nty = (size_t(Ny) + kAlphaRows - 1) / kAlphaRows;
        if ((st = dalloc(&apart, sizeof(double) * (ntx * Ny + nty * Nx + ntx * nty)))) return st;
    }
    if ((st = dalloc(&lists, sizeof(uint32_t) * 12 * nl))) return st;
    if ((st = dalloc(&FX, pb)) || (st = dalloc(&FY, pb))) return st;
    if ((st = dalloc(&pt, sizeof(double) * 4096))) return st;
    if ((st = dalloc(&sc, sizeof(Scal2)))) return st;
    hipMemset(Q, 0, pb); hipMemset(Q1, 0, pb); hipMemset(Q2, 0, pb);
    // AWCM_2D ctor (AWCM_2D.h:26-41): masks false except mask0; levels = max level
    hipMemset(mt, 0, mb); hipMemset(snap, 0, mb); hipMemset(tmp, 0, mb); hipMemset(work, 0, mb);
    for (int k = 0; k < 4; k++) { hipMemset(m[k], k == 0 ? 1 : 0, mb); hipMemset(lv[k], D.maxlev, mb); }
    hipMemset(alpha, 0, sizeof(double) * 3 * (Nx + Ny));
    Scal2 s0;
    std::memset(&s0, 0, sizeof(s0));
    s0.cnt = 1;
    s0.dirty = 1;
    if (hipMemcpy(sc, &s0, sizeof(s0), hipMemcpyHostToDevice) != hipSuccess) return AWCM_EDEVICE;
    int dev = 0, ncu = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    grid_blocks = ncu * 8;
    if (const char* e = getenv("AWCM_NO_GRAPH")) use_graph = e[0] == '0';
    return hipDeviceSynchronize() == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

int Solver2D::mark_dirty() {
    Scal2 s;
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    if (hipMemcpy(&s, sc, sizeof(s), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    s.dirty = 1;
    return hipMemcpy(sc, &s, sizeof(s), hipMemcpyHostToDevice) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

int Solver2D::upload_q(const double* q) {
    const size_t n = size_t(D.PR) * (D.Ny + 2 * G);  // points in the reference layout
    if (hipMemcpyAsync(Q2, q, sizeof(double) * 4 * n, hipMemcpyHostToDevice, stream) != hipSuccess) return AWCM_EDEVICE;
    hipLaunchKernelGGL(k2_aos_to_soa, dim3(nblk(n, 256)), dim3(256), 0, stream, Q2, Q, (long long)n, D.PR, D.P, D.plane);
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    return mark_dirty();
}

int Solver2D::download_q(double* q) {
    const size_t n = size_t(D.PR) * (D.Ny + 2 * G);
    double* stage = nullptr;
    if (hipMalloc(&stage, sizeof(double) * 4 * n) != hipSuccess) return AWCM_ENOMEM;
    hipLaunchKernelGGL(k2_soa_to_aos, dim3(nblk(n, 256)), dim3(256), 0, stream, Q, stage, (long long)n, D.PR, D.P, D.plane);
    int st = hipMemcpyAsync(q, stage, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, stream) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
    if (hipStreamSynchronize(stream) != hipSuccess) st = AWCM_EDEVICE;
    hipFree(stage);
    return st;
}

int Solver2D::set_mask0(const uint8_t* m0) {
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    if (hipMemcpy2D(m[0], D.P, m0, D.PR, D.PR, D.Ny + 2 * G, hipMemcpyHostToDevice) != hipSuccess) return AWCM_EDEVICE;
    return mark_dirty();
}

int Solver2D::get_masks(uint8_t* masks, int32_t* levels) {
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    const size_t mb = size_t(D.PR) * (D.Ny + 2 * G), nl = size_t(D.Nx) * D.Ny;
    if (masks) {
        const uint8_t* src[5] = {mt, m[0], m[1], m[2], m[3]};
        for (int k = 0; k < 5; k++)
            if (hipMemcpy2D(masks + k * mb, D.PR, src[k], D.P, D.PR, D.Ny + 2 * G, hipMemcpyDeviceToHost) != hipSuccess)
                return AWCM_EDEVICE;
    }
    if (levels) {
        int32_t* wide = nullptr;
        if (hipMalloc(&wide, sizeof(int32_t) * nl) != hipSuccess) return AWCM_ENOMEM;
        int st = AWCM_OK;
        for (int k = 0; k < 4 && !st; k++) {
            hipLaunchKernelGGL(k2_widen_levels, dim3(nblk(nl, 256)), dim3(256), 0, stream, D, lv[k], wide);
            if (hipMemcpyAsync(levels + k * nl, wide, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, stream) != hipSuccess)
                st = AWCM_EDEVICE;
            if (hipStreamSynchronize(stream) != hipSuccess) st = AWCM_EDEVICE;
        }
        hipFree(wide);
        if (st) return st;
    }
    return AWCM_OK;
}

int Solver2D::set_schedule(const double* p, int n) {
    if (n < 0 || n > 4096 || (n > 0 && !p)) return AWCM_EINVAL;
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    if (n > 0 && hipMemcpy(pt, p, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) return AWCM_EDEVICE;
    Scal2 s;
    if (hipMemcpy(&s, sc, sizeof(s), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    s.n_print = n;
    return hipMemcpy(sc, &s, sizeof(s), hipMemcpyHostToDevice) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

int Solver2D::set_state(double t, int cnt, double t_final, double tol, double cfl) {
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    Scal2 s;
    if (hipMemcpy(&s, sc, sizeof(s), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    s.t = t; s.cnt = cnt; s.t_final = t_final; s.tol = tol; s.cfl = cfl;
    s.steps = 0; s.sum_active0 = 0; s.flag = 0;
    return hipMemcpy(sc, &s, sizeof(s), hipMemcpyHostToDevice) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

int Solver2D::get_state(double* t, int* cnt, int* steps, long long* updates, int* flag) {
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    Scal2 s;
    if (hipMemcpy(&s, sc, sizeof(s), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    if (t) *t = s.t;
    if (cnt) *cnt = s.cnt;
    if (steps) *steps = s.steps;
    if (updates) *updates = s.sum_active0;
    if (flag) *flag = s.flag;
    return AWCM_OK;
}

#define AWCM_LAUNCH(...)                                             \
    do {                                                             \
        hipLaunchKernelGGL(__VA_ARGS__);                             \
        if (hipGetLastError() != hipSuccess) return AWCM_EDEVICE;    \
    } while (0)

hipEvent_t Solver2D::mark() {
    if (ev_used == ev_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ev_pool.push_back(e);
    }
    hipEvent_t e = ev_pool[ev_used++];
    if (hipEventRecord(e, stream) != hipSuccess) return nullptr;
    return e;
}

void Solver2D::prof_clear() {
    for (auto& v : prof_ev) v.clear();
    ev_used = 0;
}

constexpr int kLatK = 1;   // lattice points per thread
constexpr int kCheckK = 1; // check cells per thread

// Threads per workgroup of a lattice or check pass over n points.  Every lane
// of these passes loads from its own cache lines (stencil points are m apart
// in x and m rows apart in y), so a CU spends a cycle per lane and load on
// address processing; a coarse lattice in few full workgroups serialises on
// a few CUs.  Smaller workgroups spread it over the whole chip.
constexpr int kLatMinBlock = 16;    // smallest workgroup (one wave, 16 lanes live)
constexpr int kLatMinBlocks = 512;  // workgroup count to reach before shrinking stops
static int lat_block(long long n) {
    int b = 256;
    while (b > kLatMinBlock && nblk(n, b) < kLatMinBlocks) b >>= 1;
    return b;
}

static long long lat_threads(const Dims2& D, int lev, int kind, bool two) {
    const Lat g = lat_h(D, lev, kind);
    long long n = (long long)((g.ni + kLatK - 1) / kLatK) * g.nj;
    if (two) {
        const Lat g2 = lat_h(D, lev, LAT_D2);
        n += (long long)((g2.ni + kLatK - 1) / kLatK) * g2.nj;
    }
    return n;
}

int Solver2D::set_tail(int tail_cells) {
    // coarsest levels whose cell count fits `tail_cells` run inside one workgroup
    tail = -1;
    for (int lev = 0; lev <= D.maxlev - 1; lev++) {
        const int m = 1 << (D.L - 1 - lev);
        const long long cells = (long long)((D.Nx - 1) / m) * ((D.Ny - 1) / m);
        if (cells <= tail_cells) tail = lev;
    }
    if (gexec) { hipGraphExecDestroy(gexec); gexec = nullptr; }
    if (graph) { hipGraphDestroy(graph); graph = nullptr; }
    return AWCM_OK;
}

int Solver2D::enqueue_begin(bool force, bool copy_mt, bool prepare_only, bool clear_masks) {
    const long long nseg = (long long)(D.P / 16) * (D.Ny + 2 * G);
    const long long n = std::max<long long>(nseg, 3LL * (D.Nx + D.Ny));
    AWCM_LAUNCH(k2_begin_copy, dim3(nblk(n, 256)), dim3(256), 0, stream, D, m[0], copy_mt ? mt : m[0], alpha,
                clear_masks ? m[1] : nullptr, m[2], m[3], sc, force ? 1 : 0, prepare_only ? 1 : 0);
    return AWCM_OK;
}

// CheckReconstruction on `mask`: fused per-level passes, finest first, then the
// tail levels in one workgroup.
static int check_recon(Solver2D& s, uint8_t* mask) {
    for (int lev = s.D.maxlev - 1; lev > s.tail; lev--) {
        const int m = 1 << (s.D.L - 1 - lev);
        const long long nI = (s.D.Nx - 1) / m, nJ = (s.D.Ny - 1) / m;
        const long long n = ((nI + kCheckK - 1) / kCheckK) * nJ;
        const int b = lat_block(n);
        AWCM_LAUNCH(k2_check<kCheckK>, dim3(nblk(n, b)), dim3(b), 0, s.stream, s.D, lev, mask, s.sc);
    }
    if (s.tail >= 0) AWCM_LAUNCH(k2_check_tail, dim3(1), dim3(1024), 0, s.stream, s.D, s.tail, mask, s.sc);
    return AWCM_OK;
}

#define AWCM_LAT_ON(strm, OP, lev, kind, two, Qp, mk, aux)                                                  \
    do {                                                                                                    \
        const long long n_ = lat_threads(D, lev, kind, two);                                                \
        const int b_ = lat_block(n_);                                                                       \
        AWCM_LAUNCH((k2_lat<OP, kLatK>), dim3(nblk(n_, b_)), dim3(b_), 0, strm, D, lev, W, tol_host, Qp, mk, \
                    aux, sc);                                                                               \
    } while (0)
#define AWCM_LAT(OP, lev, kind, two, Qp, mk, aux) AWCM_LAT_ON(stream, OP, lev, kind, two, Qp, mk, aux)

int Solver2D::enqueue_prepare(bool force, bool euler) {
    const long long nseg = (long long)(D.P / 16) * (D.Ny + 2 * G);
    const unsigned bseg = nblk(nseg, 256);
    // step scalars + mask_temp = mask0 (PrepareEuler only: AWCM_2D::Prepare
    // leaves mask_temp alone, AWCM_2D.h:52-79)
    // forced: a PrepareEuler of its own, not the first part of a step
    if (int st = enqueue_begin(force, euler, force, true)) return st;
    // forward transform, finest detail level first
    for (int lev = D.maxlev - 1; lev > tail; lev--) {
        AWCM_LAT(OP_FWT_D3, lev, LAT_D3, false, Q, m[0], nullptr);
        AWCM_LAT(OP_FWT_D12, lev, LAT_D1, true, Q, m[0], nullptr);
        AWCM_LAT(OP_FWT_C, lev, LAT_C, false, Q, m[0], nullptr);
    }
    if (tail >= 0)
        AWCM_LAUNCH(k2_tail<TAIL_FWT>, dim3(1), dim3(1024), 0, stream, D, tail, W, tol_host, Q, m[0], nullptr, sc);
    AWCM_LAUNCH(k2_band_snap, dim3(bseg), dim3(256), 0, stream, D, m[0], snap, sc);
    const unsigned bmask = nblk((long long)(D.P / kMaskB) * (D.Ny + 2 * G), 256);
    AWCM_LAUNCH(k2_adjacent<kMaskB>, dim3(bmask), dim3(256), 0, stream, D, snap, m[0], sc);
    if (int st = check_recon(*this, m[0])) return st;
    for (int k = 1; k <= 3; k++) {
        AWCM_LAUNCH(k2_level_add<kMaskB>, dim3(bmask), dim3(256), 0, stream, D, m[k - 1], m[k],
                    lv[k - 1], sc);
        if (int st = check_recon(*this, m[k])) return st;
    }
    if (!euler) {
        // AWCM_2D::Prepare: the inverse transform of the scalar field (plane 0)
        // and no interpolation
        if (tail >= 0)
            AWCM_LAUNCH(k2_tail<TAIL_IWT>, dim3(1), dim3(1024), 0, stream, D, tail, W, tol_host, Q, m[3], nullptr, sc);
        for (int lev = tail + 1; lev < D.maxlev; lev++) {
            AWCM_LAT(OP_IWT_C, lev, LAT_C, false, Q, m[3], nullptr);
            AWCM_LAT(OP_IWT_D12, lev, LAT_D1, true, Q, m[3], nullptr);
            AWCM_LAT(OP_IWT_D3, lev, LAT_D3, false, Q, m[3], nullptr);
        }
        return AWCM_OK;
    }
    // mask3 is final.  Interpolation of momentum and energy at new points
    // (planes 1..3) on the side stream, the inverse transform of density
    // (plane 0) here; both only read mask3 and mask_temp.  The reference builds
    // tmp = CheckReconstruction(mask3 minus mask_temp) and interpolates where tmp
    // is set and mask_temp is not (WAVELET_2D.h:789-794,813).  mask3 is closed
    // under CheckReconstruction, so the closure only ever adds points of mask3
    // and tmp minus mask_temp is exactly mask3 minus mask_temp: the work mask is
    // mask3 itself, tested against mask_temp, and the closure pass is not needed.
    if (hipEventRecord(ev_m3, stream) != hipSuccess || hipStreamWaitEvent(s_int, ev_m3, 0) != hipSuccess)
        return AWCM_EDEVICE;
    if (tail >= 0)
        AWCM_LAUNCH(k2_tail<TAIL_INT>, dim3(1), dim3(1024), 0, s_int, D, tail, W, tol_host, Q, m[3], mt, sc);
    for (int lev = tail + 1; lev < D.maxlev; lev++) {
        AWCM_LAT_ON(s_int, OP_INT_D12, lev, LAT_D1, true, Q, m[3], mt);
        AWCM_LAT_ON(s_int, OP_INT_D3, lev, LAT_D3, false, Q, m[3], mt);
    }
    if (hipEventRecord(ev_int, s_int) != hipSuccess) return AWCM_EDEVICE;
    // inverse transform, coarsest level first (mask_temp, the previous mask0,
    // tells it where an inactive point may hold a non-zero rho)
    if (tail >= 0)
        AWCM_LAUNCH(k2_tail<TAIL_IWT>, dim3(1), dim3(1024), 0, stream, D, tail, W, tol_host, Q, m[3], nullptr, sc);
    for (int lev = tail + 1; lev < D.maxlev; lev++) {
        AWCM_LAT(OP_IWT_C, lev, LAT_C, false, Q, m[3], mt);
        AWCM_LAT(OP_IWT_D12, lev, LAT_D1, true, Q, m[3], mt);
        AWCM_LAT(OP_IWT_D3, lev, LAT_D3, false, Q, m[3], mt);
    }
    // join the side stream
    if (hipStreamWaitEvent(stream, ev_int, 0) != hipSuccess) return AWCM_EDEVICE;
    return AWCM_OK;
}

// WAVELET_2D::InterpolateAllGridEuler (WAVELET_2D.h:912-1044) on the device
// state: every finest-grid point not in mask_from gets its coarse-to-fine
// prediction, all four components, level by level (d1 and d2 of a level read
// only the level's coarse points, so they share a launch; d3 reads them).
// mask_from is in the reference layout; it is held in `tmp`.  The step
// scalars are kept, except that Q is now non-zero outside mask0.
int Solver2D::interpolate_all(const uint8_t* mask_from) {
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    if (hipMemcpy2D(tmp, D.P, mask_from, D.PR, D.PR, D.Ny + 2 * G, hipMemcpyHostToDevice) != hipSuccess)
        return AWCM_EDEVICE;
    Scal2 s0;
    if (hipMemcpy(&s0, sc, sizeof(s0), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    Scal2 s1 = s0;
    s1.skip = 0;  // the passes test `skip`, which a finished run leaves set
    if (hipMemcpy(sc, &s1, sizeof(s1), hipMemcpyHostToDevice) != hipSuccess) return AWCM_EDEVICE;
    int st = AWCM_OK;
    auto passes = [&]() -> int {
        for (int lev = 0; lev < D.maxlev; lev++) {
            AWCM_LAT(OP_IAG_D12, lev, LAT_D1, true, Q, tmp, tmp);
            AWCM_LAT(OP_IAG_D3, lev, LAT_D3, false, Q, tmp, tmp);
        }
        return AWCM_OK;
    };
    st = passes();
    if (hipStreamSynchronize(stream) != hipSuccess && !st) st = AWCM_EDEVICE;
    s0.dirty = 1;
    if (hipMemcpy(sc, &s0, sizeof(s0), hipMemcpyHostToDevice) != hipSuccess && !st) st = AWCM_EDEVICE;
    return st;
}

// WAVELET_2D::Interpolate (WAVELET_2D.h:360-473, the scalar field in plane 0)
// or InterpolateEuler(from, to) (WAVELET_2D.h:784-907, planes 1..3) on the
// device state.  The work mask is CheckReconstruction of (to minus from's
// finest-grid points), built in `work`; from is held in `tmp`.  Both masks are
// in the reference layout.  As interpolate_all, the step scalars are kept and
// the next residue pass is a full one.
int Solver2D::interpolate(const uint8_t* mask_from, const uint8_t* mask_to, bool euler) {
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    if (hipMemcpy2D(tmp, D.P, mask_from, D.PR, D.PR, D.Ny + 2 * G, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy2D(work, D.P, mask_to, D.PR, D.PR, D.Ny + 2 * G, hipMemcpyHostToDevice) != hipSuccess)
        return AWCM_EDEVICE;
    Scal2 s0;
    if (hipMemcpy(&s0, sc, sizeof(s0), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    Scal2 s1 = s0;
    s1.skip = 0;  // the check and lattice passes test `skip`
    if (hipMemcpy(sc, &s1, sizeof(s1), hipMemcpyHostToDevice) != hipSuccess) return AWCM_EDEVICE;
    auto passes = [&]() -> int {
        AWCM_LAUNCH(k2_to_minus_from, dim3(nblk(D.plane, 256)), dim3(256), 0, stream, D, tmp, work, work);
        if (int st = check_recon(*this, work)) return st;
        for (int lev = 0; lev < D.maxlev; lev++) {
            if (euler) {
                AWCM_LAT(OP_INT_D12, lev, LAT_D1, true, Q, work, tmp);
                AWCM_LAT(OP_INT_D3, lev, LAT_D3, false, Q, work, tmp);
            } else {
                AWCM_LAT(OP_SINT_D12, lev, LAT_D1, true, Q, work, tmp);
                AWCM_LAT(OP_SINT_D3, lev, LAT_D3, false, Q, work, tmp);
            }
        }
        return AWCM_OK;
    };
    int st = passes();
    if (hipStreamSynchronize(stream) != hipSuccess && !st) st = AWCM_EDEVICE;
    s0.dirty = 1;
    if (hipMemcpy(sc, &s0, sizeof(s0), hipMemcpyHostToDevice) != hipSuccess && !st) st = AWCM_EDEVICE;
    return st;
}

// The scalar field of AWCM_2D::Prepare / WAVELET_2D::Interpolate: plane 0 of Q,
// (Nx+6) x (Ny+6) doubles in the reference layout on the host.
int Solver2D::upload_scalar(const double* q) {
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    if (hipMemcpy2D(Q, sizeof(double) * D.P, q, sizeof(double) * D.PR, sizeof(double) * D.PR, D.Ny + 2 * G,
                    hipMemcpyHostToDevice) != hipSuccess)
        return AWCM_EDEVICE;
    return mark_dirty();
}
int Solver2D::download_scalar(double* q) {
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    return hipMemcpy2D(q, sizeof(double) * D.PR, Q, sizeof(double) * D.P, sizeof(double) * D.PR, D.Ny + 2 * G,
                       hipMemcpyDeviceToHost) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

// AWCM_MANAGER_2D's static mask algebra on a caller mask (reference layout,
// in/out) held in `tmp`, with the caller's levels (int32 without ghosts, in
// `work` as bytes in the masks' layout) where the operation reads or writes
// them.  The step scalars are kept; the error flag of the operation is
// returned as AWCM_ERANGE.
int Solver2D::mask_op(int op, uint8_t* mask, int32_t* levels) {
    const long long nl = (long long)D.Nx * D.Ny;
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    if (hipMemcpy2D(tmp, D.P, mask, D.PR, D.PR, D.Ny + 2 * G, hipMemcpyHostToDevice) != hipSuccess)
        return AWCM_EDEVICE;
    if (levels) {
        std::vector<uint8_t> lb(size_t(D.plane), uint8_t(D.maxlev));
        for (long long t = 0; t < nl; t++) {
            const int v = levels[t];
            if (v < 0 || v > D.maxlev) return AWCM_EINVAL;
            lb[size_t((t % D.Nx) + G + (long long)D.P * (t / D.Nx + G))] = uint8_t(v);
        }
        if (hipMemcpy(work, lb.data(), lb.size(), hipMemcpyHostToDevice) != hipSuccess) return AWCM_EDEVICE;
    }
    Scal2 s0;
    if (hipMemcpy(&s0, sc, sizeof(s0), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    Scal2 s1 = s0;
    s1.skip = 0;  // the mask passes test `skip`, which a finished run leaves set
    s1.flag = 0;
    if (hipMemcpy(sc, &s1, sizeof(s1), hipMemcpyHostToDevice) != hipSuccess) return AWCM_EDEVICE;
    int st = AWCM_OK;
    const long long ninner = (long long)(D.Nx - 2) * (D.Ny - 2);
    if (op == MASK_OP_ADJACENT || op == MASK_OP_ADD_POINT) {
        if (hipMemcpyAsync(snap, tmp, size_t(D.plane), hipMemcpyDeviceToDevice, stream) != hipSuccess) st = AWCM_EDEVICE;
        if (!st && op == MASK_OP_ADJACENT) {
            const long long n = (long long)(D.P / kMaskB) * (D.Ny + 2 * G);
            AWCM_LAUNCH((k2_adjacent<kMaskB>), dim3(nblk(n, 256)), dim3(256), 0, stream, D, snap, tmp, sc);
        } else if (!st) {
            AWCM_LAUNCH(k2_add_point, dim3(nblk(ninner, 256)), dim3(256), 0, stream, D, snap, tmp, work, sc);
        }
    } else if (op == MASK_OP_CHECK) {
        st = check_recon(*this, tmp);
    } else if (op == MASK_OP_LEVEL) {
        AWCM_LAUNCH(k2_compute_level, dim3(nblk(nl, 256)), dim3(256), 0, stream, D, tmp, work, sc);
    } else {
        st = AWCM_EINVAL;
    }
    if (hipStreamSynchronize(stream) != hipSuccess && !st) st = AWCM_EDEVICE;
    Scal2 s2;
    if (!st && hipMemcpy(&s2, sc, sizeof(s2), hipMemcpyDeviceToHost) != hipSuccess) st = AWCM_EDEVICE;
    if (hipMemcpy(sc, &s0, sizeof(s0), hipMemcpyHostToDevice) != hipSuccess && !st) st = AWCM_EDEVICE;
    if (st) return st;
    if (s2.flag & 1) return AWCM_ERANGE;
    if (op == MASK_OP_LEVEL) {
        std::vector<uint8_t> lb(size_t(D.plane));
        if (hipMemcpy(lb.data(), work, lb.size(), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
        for (long long t = 0; t < nl; t++) levels[t] = lb[size_t((t % D.Nx) + G + (long long)D.P * (t / D.Nx + G))];
        return AWCM_OK;
    }
    return hipMemcpy2D(mask, D.PR, tmp, D.P, D.PR, D.Ny + 2 * G, hipMemcpyDeviceToHost) == hipSuccess ? AWCM_OK
                                                                                                    : AWCM_EDEVICE;
}

// RemoveResidue / RemoveResidueEuler on a caller field (AoS, reference layout,
// ncomp components), staged in Q2; the mask is held in `tmp`.
int Solver2D::remove_residue(double* q, int ncomp, const uint8_t* mask) {
    const size_t n = size_t(D.PR) * (D.Ny + 2 * G) * size_t(ncomp);
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    if (hipMemcpy2D(tmp, D.P, mask, D.PR, D.PR, D.Ny + 2 * G, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(Q2, q, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
        return AWCM_EDEVICE;
    const long long ninner = (long long)(D.Nx - 2) * (D.Ny - 2);
    hipLaunchKernelGGL(k2_residue_aos, dim3(nblk(ninner, 256)), dim3(256), 0, stream, D, Q2, ncomp, tmp);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    return hipMemcpy(q, Q2, sizeof(double) * n, hipMemcpyDeviceToHost) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

int Solver2D::enqueue_step() {
    hipEvent_t e0 = prof ? mark() : nullptr;
    if (int st = enqueue_prepare(false)) return st;
    AWCM_LAUNCH(k2_compact, dim3(nblk((long long)(D.P / 16) * (D.Ny + 2 * G), kCompactThreads)), dim3(kCompactThreads), 0,
                stream, D, m[0], m[1], m[2], lv[0], lv[1], lv[2], lists, sc);
    if (prof) prof_ev[2].push_back({e0, mark()});
    const dim3 ag(nblk(D.Nx, 64), nblk(D.Ny, kAlphaRows)), ab(64, 4);
    AlphaPart ap;
    ap.px = apart;
    ap.py = apart + size_t(ag.x) * D.Ny;
    ap.pv = ap.py + size_t(ag.y) * D.Nx;
    ap.ntx = int(ag.x);
    ap.nty = int(ag.y);
    const unsigned nbc = nblk(2 * G * (D.Nx + D.Ny), 256), nred = nblk(D.Nx + D.Ny, 256);
    const unsigned gw = grid_blocks * 2, gc = grid_blocks;
    hipEvent_t pe = nullptr;
    // stage s reads Qs (ghosts filled), uses alpha set s and the lists of mask k
    struct StageArgs { int k; const double* Qa; const double* Qs; double* Qo; };
    const StageArgs sa[3] = {{2, Q, Q, Q2}, {1, Q, Q2, Q1}, {0, Q1, Q1, Q}};
    for (int s = 0; s < 3; s++) {
        const StageArgs& g = sa[s];
        // alpha partials over the stage's outer mask (interior points only),
        // then the ghost fill of Qs, which also reduces the partials (and, in
        // stage 1, turns the CFL maximum into dt)
        // (stages 2 and 3 take their alphas and ghosts from the previous
        // stage's update)
        if (s == 0) {
            AWCM_LAUNCH(k2_alpha, ag, ab, 0, stream, D, m[3], g.Qs, ap, gamma, sc);
            AWCM_LAUNCH(k2_bc, dim3(nbc + nred), dim3(256), 0, stream, D, const_cast<double*>(g.Qs), bc_kind, x0, dx,
                        1, pt, ap, ax(s), ay(s), int(nbc), ax(1), sc);
        }
        if (prof) pe = mark();
        AWCM_LAUNCH(k2_stage_faces, dim3(gw), dim3(128), 0, stream, D, list(g.k, 0), list(g.k, 2), list(g.k, 3), g.k,
                    g.Qs, FX, FY, ax(s), ay(s), sc);
        if (prof) prof_ev[0].push_back({pe, mark()});
        if (prof) pe = mark();
        if (s == 0)
            AWCM_LAUNCH(k2_stage_update<1>, dim3(gc), dim3(256), 0, stream, D, list(g.k, 0), list(g.k, 1), g.k, lv[g.k],
                        g.Qa, g.Qs, g.Qo, FX, FY, gamma, dx, ax(1), ay(1), bc_kind, x0, sc);
        else if (s == 1)
            AWCM_LAUNCH(k2_stage_update<2>, dim3(gc), dim3(256), 0, stream, D, list(g.k, 0), list(g.k, 1), g.k, lv[g.k],
                        g.Qa, g.Qs, g.Qo, FX, FY, gamma, dx, ax(2), ay(2), bc_kind, x0, sc);
        else
            AWCM_LAUNCH(k2_stage_update<3>, dim3(gc), dim3(256), 0, stream, D, list(g.k, 0), list(g.k, 1), g.k, lv[g.k],
                        g.Qa, g.Qs, g.Qo, FX, FY, gamma, dx, nullptr, nullptr, -1, x0, sc);
        if (prof) prof_ev[1].push_back({pe, mark()});
    }
    AWCM_LAUNCH(k2_residue, dim3(nblk(D.Nx, 256), D.Ny), dim3(256), 0, stream, D, m[0], m[3], mt, Q, 1, sc);
    return AWCM_OK;
}

int Solver2D::launch_step() {
    if (!use_graph || prof) return enqueue_step();
    if (!gexec || graph_tol != tol_host) {
        if (gexec) { hipGraphExecDestroy(gexec); gexec = nullptr; }
        if (graph) { hipGraphDestroy(graph); graph = nullptr; }
        if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return AWCM_EDEVICE;
        const int st = enqueue_step();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(stream, &g);
        if (st) { if (g) hipGraphDestroy(g); return st; }
        if (e != hipSuccess || !g) return AWCM_EDEVICE;
        graph = g;
        if (hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess) { gexec = nullptr; return AWCM_EDEVICE; }
        graph_tol = tol_host;
    }
    return hipGraphLaunch(gexec, stream) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

// n_steps iterations of Wavelet/main_2d_advect.cpp:46-199 on plane 0 of Q:
// AWCM_2D::Prepare, the periodic ghost fill, the three stages on mask2,
// mask1, mask0 (reading Q, A2, A1), the combine and PostProcess.
int Solver2D::advect(double tol, double dt, int n_steps) {
    if (n_steps < 0) return AWCM_EINVAL;
    const size_t pl = sizeof(double) * size_t(D.plane);
    if (!adv) {
        if (int st = dalloc(&adv, 3 * pl)) return st;
        if (hipMemset(adv, 0, 3 * pl) != hipSuccess) return AWCM_EDEVICE;
    }
    double* A2 = adv;
    double* A1 = adv + D.plane;
    double* A0 = adv + 2 * D.plane;
    Scal2 s0;
    if (hipStreamSynchronize(stream) != hipSuccess || hipMemcpy(&s0, sc, sizeof(s0), hipMemcpyDeviceToHost) != hipSuccess)
        return AWCM_EDEVICE;
    s0.flag = 0;
    if (hipMemcpy(sc, &s0, sizeof(s0), hipMemcpyHostToDevice) != hipSuccess) return AWCM_EDEVICE;
    tol_host = tol;
    const dim3 pg(nblk(D.Nx, 256), D.Ny), pb(256);
    const unsigned nper = nblk(2 * G * (D.Nx + D.Ny), 256);
    for (int k = 0; k < n_steps; k++) {
        if (int st = enqueue_prepare(true, false)) return st;
        AWCM_LAUNCH(k2_periodic, dim3(nper), dim3(256), 0, stream, D, Q);
        AWCM_LAUNCH(k2_advect_stage<1>, pg, pb, 0, stream, D, m[2], lv[2], Q, Q, A2, dt, dx, sc);
        AWCM_LAUNCH(k2_periodic, dim3(nper), dim3(256), 0, stream, D, A2);
        AWCM_LAUNCH(k2_advect_stage<2>, pg, pb, 0, stream, D, m[1], lv[1], Q, A2, A1, dt, dx, sc);
        AWCM_LAUNCH(k2_periodic, dim3(nper), dim3(256), 0, stream, D, A1);
        AWCM_LAUNCH(k2_advect_stage<3>, pg, pb, 0, stream, D, m[0], lv[0], Q, A1, A0, dt, dx, sc);
        AWCM_LAUNCH(k2_advect_combine, pg, pb, 0, stream, D, m[0], Q, A0);
    }
    if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
    Scal2 s1;
    if (hipMemcpy(&s1, sc, sizeof(s1), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    return (s1.flag & 1) ? AWCM_ERANGE : AWCM_OK;
}

int Solver2D::enqueue_post() {
    if (int st = enqueue_begin(true, false, false, false)) return st;
    AWCM_LAUNCH(k2_residue, dim3(nblk(D.Nx, 256), D.Ny), dim3(256), 0, stream, D, m[0], m[3], mt, Q, 0, sc);
    return AWCM_OK;
}

}  // namespace awcm
