// 1-D AWCM compressible-Euler solver on gfx950.
//
// One workgroup owns one independent solve (a "tube").  Its five masks and four
// level fields live in LDS for the whole launch (in HBM, workgroup-private, for
// instances too long for one workgroup's LDS); the conserved state Q, the RK
// copies and the face-flux buffer live in HBM (planes of Nx+6 doubles, SoA) and
// stay L2-resident at these sizes.  A launch of a single instance copies them
// into LDS beside the masks for the launch when they fit (kModeLdsState), or
// Q alone (kModeLdsQ); a batched launch of one workgroup per CU does the same
// instance by instance (kModeLdsAuto).  A launch runs the reference's time loop
// (Wavelet/main_1d_euler.cpp:61-260, with the WENO-Z flux pair of the C1
// configuration or, FLUX = kRecWeno5, the shipped driver's WENO5-JS ComputedF)
// to t_final or a step budget without returning to the host,
// so the plain solve is a single kernel and the batched form is the same kernel
// over a work queue of instances (largest first), one workgroup each.
//
// Bit-exactness: every arithmetic expression follows the reference's operation
// order (see awcm/physics.h); the only restructuring is that each WENO-Z face
// flux is evaluated once and shared by the two points beside it -- the
// reference evaluates the same function on the same six states twice
// (main_1d_euler.cpp:109 via WENO.h:176-184), so results are identical.

#include "awcm_capi.h"
#include "solver1d_dev.h"

namespace awcm {

// the C ABI's flux kinds name the reconstructions of awcm/physics.h
static_assert(AWCM_FLUX_WENOZ == kRecWenoZ && AWCM_FLUX_WENO5 == kRecWeno5, "flux kinds");

// mode 0: PrepareEuler only; mode 2: PostProcessEuler only; mode 1: time
// steps until t >= t_final or max_steps.  The SCALAR instantiation (kept apart
// so the time-step kernel is unchanged) runs mode 3, the scalar Prepare (plane
// 0), and the interpolations from mask_temp's slot onto mask3's slot (the host
// puts the two masks there and restores them): mode 4 the scalar one (plane
// 0), mode 5 InterpolateEuler(from, to) (planes 1..2), mode 6 the one-mask
// InterpolateEuler (all three planes, WAVELET_1D.h:433-473); modes 7..10
// AddAdjacentZone, CheckReconstruction, AddPointForDifferentiation and
// ComputeLevel (AWCM_MANAGER_1D.h:14-133) on mask3's and level3's slots.
// steps of this launch.  Instances are taken from `order` through the atomic
// work counter *qctr (the host zeroes it).  The masks and levels of the
// instance in hand occupy 11 * nxp bytes (nxp >= max(Nx) + 6 over the
// instances): dynamic LDS, or with kHbm -- for instances whose masks do not
// fit one workgroup's LDS -- this workgroup's slice of `hbm` (gridDim.x slices
// of 11 * nxp bytes; workgroup-local, ordered by the same barriers).
template <bool kHbm, int FLUX, bool SCALAR>
__global__ void __launch_bounds__(1024) k1_solve(Inst1D* insts, int n_inst, const int* order, int* qctr, int mode,
                                                 int max_steps, int nxp, uint8_t* hbm) {
    AWCM_DYNAMIC_LDS __align__(16) uint8_t lds_masks[];
    uint8_t* const sm = kHbm ? hbm + size_t(blockIdx.x) * 11 * size_t(nxp) : lds_masks;
    const int lflags = mode & (kModeLdsState | kModeLdsQ | kModeLdsAuto);
    const size_t lbudget = size_t(unsigned(mode) >> 16) * 16;
    mode &= kModeBase;
    __shared__ double W[4][5];
    __shared__ double red[16];
    __shared__ int s_idx, s_flag;
    __shared__ LoopState st;
    if (threadIdx.x == 0) wavelet_weights(W);
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_idx = atomicAdd(qctr, 1);
        __syncthreads();
        const int idx = s_idx;
        if (idx >= n_inst) break;
        Inst1D& in = insts[order[idx]];
        Ctx1 c;
        c.g.L = in.L; c.g.maxlev = in.L - 1; c.g.Nx = in.Nx; c.g.nxc = in.nxc;
        c.gamma = in.gamma; c.dx = in.dx;
        c.NP = in.Nx + 2 * G;
        c.Q = in.Q; c.Qc = in.Qc; c.Qt = in.Qt; c.F = in.F;
        const int nq = 3 * c.NP, nf = 3 * (in.Nx + 1);
        // where this instance's state lives for the launch (planes ps apart)
        bool lstate = false, lq = false;
        int ps = nxp;
        if (!kHbm && (lflags & kModeLdsAuto)) {
            ps = c.NP;
            lstate = lds_state_offset(nxp) + lds_state_bytes(ps) <= lbudget;
            lq = !lstate && lds_state_offset(nxp) + lds_q_bytes(ps) <= lbudget;
        } else if (!kHbm) {
            lstate = (lflags & kModeLdsState) != 0;
            lq = !lstate && (lflags & kModeLdsQ) != 0;
        }
        if (lstate) {
            double* ls = reinterpret_cast<double*>(lds_masks + lds_state_offset(nxp));
            for (int k = threadIdx.x; k < nq; k += blockDim.x) {
                ls[k] = in.Q[k];
                ls[3 * ps + k] = in.Qc[k];
                ls[6 * ps + k] = in.Qt[k];
            }
            for (int k = threadIdx.x; k < nf; k += blockDim.x) ls[9 * ps + k] = in.F[k];
            c.Q = ls; c.Qc = ls + 3 * ps; c.Qt = ls + 6 * ps; c.F = ls + 9 * ps;
        } else if (lq) {
            double* ls = reinterpret_cast<double*>(lds_masks + lds_state_offset(nxp));
            for (int k = threadIdx.x; k < nq; k += blockDim.x) ls[k] = in.Q[k];
            c.Q = ls;
        }
        c.mt = sm;
        for (int k = 0; k < 4; k++) c.m[k] = sm + (k + 1) * nxp;
        c.snap = sm + 5 * nxp;
        c.tmp = sm + 6 * nxp;
        for (int k = 0; k < 4; k++) c.lv[k] = sm + (7 + k) * nxp;
        c.W = W; c.red = red; c.flag = &s_flag;
        c.rec = (mode == 1 && in.tape_m) ? &in : nullptr;
        if (threadIdx.x == 0) s_flag = 0;
        // load persistent masks (ghosts included) and levels
        for (int k = 0; k < 5; k++)
            for (int i = threadIdx.x; i < c.NP; i += blockDim.x) sm[k * nxp + i] = in.masks[k * c.NP + i];
        for (int k = 0; k < 4; k++)
            for (int i = threadIdx.x; i < in.Nx; i += blockDim.x) c.lv[k][i] = in.levels[k * in.Nx + i];
        __syncthreads();
        if (threadIdx.x == 0) {
            st.t = in.t; st.tol = in.tol; st.cfl = in.cfl; st.t_final = in.t_final;
            st.sum_active0 = 0; st.cnt = in.cnt; st.steps = 0; st.n_print = in.n_print; st.bad = 0;
        }
        __syncthreads();
        if constexpr (SCALAR) {
            if (mode == 3) prepare_scalar(c, st.tol);
            else if (mode == 4) interp<0, 1>(c, c.mt, c.m[3]);
            else if (mode == 5) interp<1, 3>(c, c.mt, c.m[3]);
            else if (mode == 6) interp<0, 3>(c, c.mt, c.m[3]);
            // AWCM_MANAGER_1D's statics on the mask in mask3's slot (and the
            // level field in level3's), which the host fills and restores
            else if (mode == 7) adjacent(c, c.m[3]);
            else if (mode == 8) check_recon(c, c.m[3]);
            else if (mode == 9) addpoint(c, c.m[3], c.lv[3]);
            else if (mode == 10) level(c, c.lv[3], c.m[3]);
        } else if (mode == 0) {
            prepare(c, st.tol);
        } else if (mode == 2) {
            post(c);
        } else {
            const int budget = c.rec ? min(max_steps, in.tape_cap - in.tape_n) : max_steps;
            for (int s = 0; s < budget; s++) {
                if (!(st.t < st.t_final) || st.bad) break;
                step<FLUX>(c, st, in.print_times);
            }
        }
        __syncthreads();
        if (lstate) {
            for (int k = threadIdx.x; k < nq; k += blockDim.x) {
                in.Q[k] = c.Q[k];
                in.Qc[k] = c.Qc[k];
                in.Qt[k] = c.Qt[k];
            }
            for (int k = threadIdx.x; k < nf; k += blockDim.x) in.F[k] = c.F[k];
        } else if (lq) {
            for (int k = threadIdx.x; k < nq; k += blockDim.x) in.Q[k] = c.Q[k];
        }
        for (int k = 0; k < 5; k++)
            for (int i = threadIdx.x; i < c.NP; i += blockDim.x) in.masks[k * c.NP + i] = sm[k * nxp + i];
        for (int k = 0; k < 4; k++)
            for (int i = threadIdx.x; i < in.Nx; i += blockDim.x) in.levels[k * in.Nx + i] = c.lv[k][i];
        if (threadIdx.x == 0) {
            if (s_flag) in.status = -5;
            if (st.bad) in.status = -4;
            in.t = st.t; in.cnt = st.cnt; in.steps += st.steps; in.sum_active0 += st.sum_active0;
            if (c.rec) in.tape_n += st.steps;
        }
    }
}

template __global__ void k1_solve<false, kRecWenoZ, false>(Inst1D*, int, const int*, int*, int, int, int, uint8_t*);
template __global__ void k1_solve<true, kRecWenoZ, false>(Inst1D*, int, const int*, int*, int, int, int, uint8_t*);
template __global__ void k1_solve<false, kRecWeno5, false>(Inst1D*, int, const int*, int*, int, int, int, uint8_t*);
template __global__ void k1_solve<true, kRecWeno5, false>(Inst1D*, int, const int*, int*, int, int, int, uint8_t*);
template __global__ void k1_solve<false, kRecWenoZ, true>(Inst1D*, int, const int*, int*, int, int, int, uint8_t*);
template __global__ void k1_solve<true, kRecWenoZ, true>(Inst1D*, int, const int*, int*, int, int, int, uint8_t*);

}  // namespace awcm
