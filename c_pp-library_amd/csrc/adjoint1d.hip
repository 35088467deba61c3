// Host side of the 1-D discrete adjoint: tape allocation and the C ABI entry
// points awcm_adjoint_record / awcm_adjoint_replay / awcm_adjoint.  The device
// code is in adjoint1d_dev.h (frozen replay and segment recompute, one
// workgroup) and adjoint1d_rev.h (the reverse sweep, grid-wide kernels).

#include "adjoint1d_rev.h"

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "awcm_capi.h"
#include "runtime.h"

namespace awcm {

Adj1DHost::~Adj1DHost() {
    if (gexec) hipGraphExecDestroy(gexec);
    if (graph) hipGraphDestroy(graph);
    for (void* p : allocs) hipFree(p);
}

int Adj1DHost::dalloc(void* pp, size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return AWCM_ENOMEM;
    allocs.push_back(p);
    *static_cast<void**>(pp) = p;
    return AWCM_OK;
}

namespace {

int nxp_of(const Inst1D& in) { return ((in.Nx + 2 * G) + 15) / 16 * 16; }

Adj1DArgs args_of(const Adj1DHost& H) {
    Adj1DArgs A;
    std::memset(&A, 0, sizeof(A));
    A.tape_m = H.tape_m;
    A.tape_s = H.tape_s;
    A.ckpt = H.ckpt;
    A.tape_stride = H.stride;
    A.n_steps = H.steps;
    A.ckpt_every = H.K;
    A.seg_steps = H.seg_steps;
    A.seg = H.seg;
    A.step = H.d_step;
    return A;
}

// k1_adjoint keeps the masks in LDS only: an instance whose 11 * nxp bytes do
// not fit one workgroup's LDS next to the kernel's static LDS is refused.  When
// the work state fits beside them as well, it lives there too (*lq set, *lds
// the dynamic LDS bytes).
int adjoint_fits(int nxp, int* lq = nullptr, size_t* lds = nullptr) {
    int dev = 0, cap = 0;
    hipFuncAttributes fa;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cap, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k1_adjoint)) != hipSuccess)
        return AWCM_EDEVICE;
    if (size_t(11) * nxp + fa.sharedSizeBytes > size_t(cap)) return AWCM_EINVAL;
    const size_t with_q = lds_state_offset(nxp) + lds_q_bytes(nxp);
    const bool q = with_q + fa.sharedSizeBytes <= size_t(cap);
    if (lq) *lq = q ? 1 : 0;
    if (lds) *lds = q ? with_q : size_t(11) * nxp;
    return AWCM_OK;
}

int launch_adjoint(Batch1D& B, const Adj1DHost& H, int mode) {
    const int nxp = nxp_of(B.h[0]);
    int lq = 0;
    size_t lds = 0;
    if (int st = adjoint_fits(nxp, &lq, &lds)) return st;
    hipLaunchKernelGGL(k1_adjoint, dim3(1), dim3(1024), lds, B.stream, B.d_inst, args_of(H), mode, nxp, lq);
    if (hipGetLastError() != hipSuccess) return AWCM_EDEVICE;
    int st = B.pull();
    if (!st && B.h[0].status) st = B.h[0].status == -5 ? AWCM_ERANGE : AWCM_EDEVICE;
    return st;
}

}  // namespace
}  // namespace awcm

using awcm::Adj1DHost;
using awcm::Batch1D;

extern "C" {

int awcm_adjoint_record(awcm_solver* h, double tol, double cfl, double t_final, int max_steps, int checkpoint_every,
                        double* t, int* cnt, int* steps_done) {
    if (!h || !t || !cnt || *cnt < 0 || max_steps <= 0 || checkpoint_every <= 0 || !(cfl > 0)) return AWCM_EINVAL;
    if (h->dim != 1) return AWCM_ESTATE;
    Batch1D& B = h->b1;
    awcm::Inst1D& in = B.h[0];
    if (B.flux != AWCM_FLUX_WENOZ) return AWCM_ESTATE;  // the adjoint differentiates the WENO-Z flux pair
    delete h->adj;
    h->adj = new (std::nothrow) Adj1DHost();
    if (!h->adj) return AWCM_ENOMEM;
    Adj1DHost& H = *h->adj;
    const long long NP = in.Nx + 2 * awcm::G;
    H.stride = awcm::adj_tape_stride(in.Nx);
    H.K = checkpoint_every;
    H.cap = max_steps;
    int st = H.dalloc(&H.tape_m, size_t(H.stride) * max_steps);
    if (!st) st = H.dalloc(&H.tape_s, sizeof(double) * 4 * size_t(max_steps));
    if (!st) st = H.dalloc(&H.ckpt, sizeof(double) * 3 * NP * size_t(max_steps / checkpoint_every + 1));
    if (st) { delete h->adj; h->adj = nullptr; return st; }
    in.tol = tol; in.cfl = cfl; in.t_final = t_final; in.t = *t; in.cnt = *cnt;
    in.steps = 0; in.sum_active0 = 0; in.status = 0;
    in.tape_m = H.tape_m; in.tape_s = H.tape_s; in.ckpt = H.ckpt;
    in.tape_stride = H.stride; in.tape_n = 0; in.tape_cap = max_steps; in.ckpt_every = checkpoint_every;
    st = B.push();
    if (!st) st = B.launch(1, max_steps, 0, 1);
    if (!st) st = B.pull();
    H.steps = in.tape_n;
    in.tape_m = nullptr; in.tape_s = nullptr; in.ckpt = nullptr;
    in.tape_n = 0; in.tape_cap = 0;
    if (!st) {
        *t = in.t; *cnt = in.cnt;
        if (steps_done) *steps_done = in.steps;
        st = B.check_status();
    }
    return st;
}

int awcm_adjoint_replay(awcm_solver* h) {
    if (!h) return AWCM_EINVAL;
    if (h->dim != 1 || !h->adj) return AWCM_ESTATE;
    Batch1D& B = h->b1;
    B.h[0].status = 0;
    int st = B.push();
    if (!st) st = awcm::launch_adjoint(B, *h->adj, 0);
    return st;
}

int awcm_adjoint(awcm_solver* h, const double* dJ_dQfinal, int checkpoint_every, double* dJ_dQ0) {
    if (!h || !dJ_dQfinal || !dJ_dQ0 || checkpoint_every < 0) return AWCM_EINVAL;
    if (h->dim != 1 || !h->adj) return AWCM_ESTATE;
    Adj1DHost& H = *h->adj;
    const int K = checkpoint_every ? checkpoint_every : H.K;
    if (K % H.K) return AWCM_EINVAL;
    Batch1D& B = h->b1;
    awcm::Inst1D& in = B.h[0];
    const size_t NP = size_t(in.Nx) + 2 * awcm::G;
    const int nxp = awcm::nxp_of(in);
    int lq = 0;
    size_t lds = 0;
    int st = awcm::adjoint_fits(nxp, &lq, &lds);
    if (st) return st;
    if (!H.Qb) {
        st = H.dalloc(&H.Qb, sizeof(double) * 3 * NP);
        if (!st) st = H.dalloc(&H.Xb, sizeof(double) * 3 * NP);
        if (!st) st = H.dalloc(&H.Cb, sizeof(double) * 3 * NP);
        if (!st) st = H.dalloc(&H.Gb, sizeof(double) * 3 * NP);
        if (!st) st = H.dalloc(&H.FV, sizeof(double) * 18 * (size_t(in.Nx) + 1));
        if (!st) st = H.dalloc(&H.Qsave, sizeof(double) * 3 * NP);
        if (!st) st = H.dalloc(&H.d_step, sizeof(int));
        if (st) return st;
    }
    if (H.seg_cap < K) {
        if ((st = H.dalloc(&H.seg, sizeof(double) * 9 * NP * size_t(K)))) return st;
        H.seg_cap = K;
    }
    H.seg_steps = K;
    if (H.steps == 0) {  // nothing recorded: the gradient passes through unchanged
        for (size_t i = 0; i < NP; i++)
            for (int c = 0; c < 3; c++)
                dJ_dQ0[i * 3 + c] = (i < size_t(awcm::G) || i >= NP - awcm::G) ? 0.0 : dJ_dQfinal[i * 3 + c];
        return AWCM_OK;
    }
    // one reversed step, captured once per segment length and replayed for
    // every step: recompute the segment if the step ends it, reverse the step,
    // move the step counter back
    if (!H.gexec || H.g_K != K || H.g_seg != H.seg) {
        if (H.gexec) { hipGraphExecDestroy(H.gexec); H.gexec = nullptr; }
        if (H.graph) { hipGraphDestroy(H.graph); H.graph = nullptr; }
        awcm::AdjRev R;
        std::memset(&R, 0, sizeof(R));
        R.Nx = in.Nx; R.L = in.L; R.NP = int(NP); R.K = K;
        R.gamma = in.gamma; R.dx = in.dx;
        R.tape_m = H.tape_m; R.stride = H.stride; R.tape_s = H.tape_s; R.seg = H.seg;
        R.Qb = H.Qb; R.Xb = H.Xb; R.Cb = H.Cb; R.Gb = H.Gb; R.FV = H.FV;
        R.step = H.d_step;
        const awcm::Adj1DArgs A = awcm::args_of(H);
        hipStream_t stream = B.stream;
        if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return AWCM_EDEVICE;
        hipLaunchKernelGGL(awcm::k1_adjoint, dim3(1), dim3(1024), lds, stream, B.d_inst, A, 2, nxp, lq);
        awcm::launch_reverse_step(R, stream);
        hipLaunchKernelGGL(awcm::k1r_next, dim3(1), dim3(64), 0, stream, H.d_step);
        if (hipStreamEndCapture(stream, &H.graph) != hipSuccess) return AWCM_EDEVICE;
        if (hipGraphInstantiate(&H.gexec, H.graph, nullptr, nullptr, 0) != hipSuccess) return AWCM_EDEVICE;
        H.g_K = K;
        H.g_seg = H.seg;
    }
    std::vector<double> soa(3 * NP, 0.0);
    for (size_t i = awcm::G; i < NP - awcm::G; i++)
        for (int c = 0; c < 3; c++) soa[c * NP + i] = dJ_dQfinal[i * 3 + c];
    const int last = H.steps - 1;
    if (hipStreamSynchronize(B.stream) != hipSuccess) return AWCM_EDEVICE;
    if (hipMemcpyAsync(H.Qb, soa.data(), sizeof(double) * 3 * NP, hipMemcpyHostToDevice, B.stream) != hipSuccess ||
        hipMemcpyAsync(H.Qsave, in.Q, sizeof(double) * 3 * NP, hipMemcpyDeviceToDevice, B.stream) != hipSuccess ||
        hipMemcpyAsync(H.d_step, &last, sizeof(int), hipMemcpyHostToDevice, B.stream) != hipSuccess)
        return AWCM_EDEVICE;
    in.status = 0;
    st = B.push();
    for (int n = 0; n < H.steps && !st; n++)
        if (hipGraphLaunch(H.gexec, B.stream) != hipSuccess) st = AWCM_EDEVICE;
    if (hipMemcpyAsync(in.Q, H.Qsave, sizeof(double) * 3 * NP, hipMemcpyDeviceToDevice, B.stream) != hipSuccess && !st)
        st = AWCM_EDEVICE;
    if (!st && hipMemcpyAsync(soa.data(), H.Qb, sizeof(double) * 3 * NP, hipMemcpyDeviceToHost, B.stream) != hipSuccess)
        st = AWCM_EDEVICE;
    const int pst = B.pull();  // synchronises the stream
    if (!st) st = pst;
    if (!st && B.h[0].status) st = B.h[0].status == -5 ? AWCM_ERANGE : AWCM_EDEVICE;
    if (st) return st;
    for (size_t i = 0; i < NP; i++)
        for (int c = 0; c < 3; c++)
            dJ_dQ0[i * 3 + c] = (i < size_t(awcm::G) || i >= NP - awcm::G) ? 0.0 : soa[c * NP + i];
    return AWCM_OK;
}

}  // extern "C"
