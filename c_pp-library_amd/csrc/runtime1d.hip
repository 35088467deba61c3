// Host runtime of the 1-D solver and of the batched 1-D form behind the C ABI.
// A plain 1-D solver handle is a batch of one instance.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "awcm_capi.h"
#include "awcm_internal.h"
#include "runtime.h"

namespace awcm {

template <bool kHbm, int FLUX, bool SCALAR>
__global__ void k1_solve(Inst1D* insts, int n_inst, const int* order, int* qctr, int mode, int max_steps, int nxp,
                         uint8_t* hbm);

// The k1_solve instantiation for a mask placement and a stage flux; modes 3 and
// 4 (the scalar path) have their own.
static void launch_solve(bool hbm_masks, int flux, dim3 grid, dim3 block, size_t lds, hipStream_t stream, Inst1D* insts,
                         int n_inst, const int* order, int* qctr, int mode, int max_steps, int nxp, uint8_t* hbm) {
    if ((mode & kModeBase) >= 3) {
        if (hbm_masks)
            hipLaunchKernelGGL((k1_solve<true, AWCM_FLUX_WENOZ, true>), grid, block, lds, stream, insts, n_inst, order,
                               qctr, mode, max_steps, nxp, hbm);
        else
            hipLaunchKernelGGL((k1_solve<false, AWCM_FLUX_WENOZ, true>), grid, block, lds, stream, insts, n_inst, order,
                               qctr, mode, max_steps, nxp, hbm);
    } else if (flux == AWCM_FLUX_WENO5) {
        if (hbm_masks)
            hipLaunchKernelGGL((k1_solve<true, AWCM_FLUX_WENO5, false>), grid, block, lds, stream, insts, n_inst, order,
                               qctr, mode, max_steps, nxp, hbm);
        else
            hipLaunchKernelGGL((k1_solve<false, AWCM_FLUX_WENO5, false>), grid, block, lds, stream, insts, n_inst, order,
                               qctr, mode, max_steps, nxp, hbm);
    } else {
        if (hbm_masks)
            hipLaunchKernelGGL((k1_solve<true, AWCM_FLUX_WENOZ, false>), grid, block, lds, stream, insts, n_inst, order,
                               qctr, mode, max_steps, nxp, hbm);
        else
            hipLaunchKernelGGL((k1_solve<false, AWCM_FLUX_WENOZ, false>), grid, block, lds, stream, insts, n_inst, order,
                               qctr, mode, max_steps, nxp, hbm);
    }
}

// Work-queue counter reset as a kernel, so the batch launch can be captured
// into a graph without a memset node.
__global__ void k1_reset_counter(int* qctr) {
    if (threadIdx.x == 0) *qctr = 0;
}

namespace {

constexpr int G = 3;

size_t lds_bytes(int nxp) { return size_t(11) * nxp; }

}  // namespace

Batch1D::~Batch1D() {
    if (gexec) hipGraphExecDestroy(gexec);
    if (graph) hipGraphDestroy(graph);
    for (void* p : allocs) hipFree(p);
    if (stream) hipStreamDestroy(stream);
}

int Batch1D::init(int n_, const awcm_instance_1d* params, const double* pt, int npt) {
    n = n_;
    if (n <= 0 || !params) return AWCM_EINVAL;
    if (npt < 0 || (npt > 0 && !pt)) return AWCM_EINVAL;
    for (int k = 0; k < n; k++) {
        const awcm_instance_1d& p = params[k];
        if (p.num_levels < 2 || p.num_levels > 14 || p.nx_coarse < 6) return AWCM_EINVAL;
        if (!(p.dx > 0) || !(p.gamma > 1) || !(p.cfl > 0)) return AWCM_EINVAL;
        if ((long long)(p.nx_coarse - 1) * (1 << (p.num_levels - 1)) + 1 > 65535) return AWCM_EINVAL;
    }
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return AWCM_EDEVICE;
    {
        int dev = 0, cap = 0;
        hipFuncAttributes fa;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cap, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
            hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k1_solve<false, AWCM_FLUX_WENOZ, false>)) != hipSuccess)
            return AWCM_EDEVICE;
        lds_cap = size_t(cap) > fa.sharedSizeBytes ? size_t(cap) - fa.sharedSizeBytes : 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return AWCM_EDEVICE;
    }
    h.resize(n);
    max_nx = 0;
    for (int k = 0; k < n; k++) {
        const awcm_instance_1d& p = params[k];
        Inst1D& in = h[k];
        std::memset(&in, 0, sizeof(in));
        in.L = p.num_levels;
        in.nxc = p.nx_coarse;
        in.Nx = (p.nx_coarse - 1) * (1 << (p.num_levels - 1)) + 1;
        in.gamma = p.gamma;
        in.dx = p.dx;
        in.tol = p.tol;
        in.cfl = p.cfl;
        in.t_final = p.t_final;
        in.cnt = 1;
        max_nx = std::max(max_nx, in.Nx);
    }
    // device buffers
    int st = dalloc(&d_pt, sizeof(double) * std::max(npt, 1));
    if (st) return st;
    if (npt > 0 && hipMemcpy(d_pt, pt, sizeof(double) * npt, hipMemcpyHostToDevice) != hipSuccess) return AWCM_EDEVICE;
    n_print = npt;
    for (int k = 0; k < n; k++) {
        Inst1D& in = h[k];
        const int NP = in.Nx + 2 * G;
        in.n_print = npt;
        in.print_times = d_pt;
        if ((st = dalloc(&in.Q, sizeof(double) * 3 * NP))) return st;
        if ((st = dalloc(&in.Qc, sizeof(double) * 3 * NP))) return st;
        if ((st = dalloc(&in.Qt, sizeof(double) * 3 * NP))) return st;
        if ((st = dalloc(&in.F, sizeof(double) * 3 * (in.Nx + 1)))) return st;
        if ((st = dalloc(&in.masks, size_t(5) * NP))) return st;
        if ((st = dalloc(&in.levels, size_t(4) * in.Nx))) return st;
        hipMemset(in.Q, 0, sizeof(double) * 3 * NP);
        hipMemset(in.Qc, 0, sizeof(double) * 3 * NP);
        hipMemset(in.Qt, 0, sizeof(double) * 3 * NP);
        // AWCM_1D ctor (AWCM_1D.h:23-34): all masks false except mask0, levels 0
        hipMemset(in.masks, 0, size_t(5) * NP);
        hipMemset(in.masks + NP, 1, NP);
        hipMemset(in.levels, 0, size_t(4) * in.Nx);
    }
    if ((st = dalloc(&d_inst, sizeof(Inst1D) * n))) return st;
    if ((st = dalloc(&d_order, sizeof(int) * n))) return st;
    if ((st = dalloc(&d_qctr, sizeof(int)))) return st;
    // largest instance first (longest-processing-time order for the work queue)
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return h[a].Nx > h[b].Nx; });
    if (hipMemcpy(d_order, order.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess) return AWCM_EDEVICE;
    if (hipDeviceSynchronize() != hipSuccess) return AWCM_EDEVICE;
    threads = max_nx >= 1024 ? 1024 : 256;
    workgroups = 0;
    return AWCM_OK;
}

int Batch1D::dalloc(void* pp, size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return AWCM_ENOMEM;
    allocs.push_back(p);
    *static_cast<void**>(pp) = p;
    return AWCM_OK;
}

int Batch1D::push() {
    return hipMemcpyAsync(d_inst, h.data(), sizeof(Inst1D) * n, hipMemcpyHostToDevice, stream) == hipSuccess
               ? AWCM_OK : AWCM_EDEVICE;
}

int Batch1D::pull() {
    if (hipMemcpyAsync(h.data(), d_inst, sizeof(Inst1D) * n, hipMemcpyDeviceToHost, stream) != hipSuccess)
        return AWCM_EDEVICE;
    return hipStreamSynchronize(stream) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

// Where a launch of wgs workgroups keeps its masks: LDS (*hbm = nullptr) when
// 11 * nxp bytes fit, else a slice per workgroup of an HBM buffer.  Never
// called during graph capture (it may allocate).
int Batch1D::mask_space(int wgs, int nxp, uint8_t** hbm) {
    *hbm = nullptr;
    if (lds_bytes(nxp) <= lds_cap) return AWCM_OK;
    const size_t need = lds_bytes(nxp) * size_t(wgs);
    if (need > hbm_masks_bytes) {
        if (hipStreamSynchronize(stream) != hipSuccess) return AWCM_EDEVICE;
        if (d_hbm_masks) {
            allocs.erase(std::find(allocs.begin(), allocs.end(), static_cast<void*>(d_hbm_masks)));
            hipFree(d_hbm_masks);
            d_hbm_masks = nullptr;
            hbm_masks_bytes = 0;
        }
        if (int st = dalloc(&d_hbm_masks, need)) return st;
        hbm_masks_bytes = need;
    }
    *hbm = d_hbm_masks;
    return AWCM_OK;
}

// A launch of at most one workgroup per CU: its workgroups take the whole LDS
// budget (a CU holds one of them either way), and each instance keeps its
// state or Q in the part its masks leave (kModeLdsAuto).  Otherwise the
// masks alone.
int Batch1D::lds_auto(int mode, int wgs, size_t* lds) const {
    if (!lds_state || wgs > ncu || lds_cap <= *lds) return mode;
    *lds = lds_cap / 16 * 16;
    return mode | kModeLdsAuto | int((*lds / 16) << 16);
}

int Batch1D::launch(int mode, int max_steps, int first, int count) {
    const int nxp = ((max_nx + 2 * G) + 15) / 16 * 16;
    const int wgs = workgroups > 0 ? std::min(workgroups, count) : count;
    uint8_t* hbm = nullptr;
    if (int st = mask_space(wgs, nxp, &hbm)) return st;
    if (hipMemsetAsync(d_qctr, 0, sizeof(int), stream) != hipSuccess) return AWCM_EDEVICE;
    // one instance whose state fits beside its masks: the state in LDS too
    // (its every access is then an LDS access instead of an L2 round trip)
    size_t lds = hbm ? 0 : lds_bytes(nxp);
    if (!hbm && count == 1 && lds_state && lds_state_offset(nxp) + lds_state_bytes(nxp) <= lds_cap) {
        mode |= kModeLdsState;
        lds = lds_state_offset(nxp) + lds_state_bytes(nxp);
    } else if (!hbm && count == 1 && lds_state && lds_state_offset(nxp) + lds_q_bytes(nxp) <= lds_cap) {
        // too long for all of it: Q alone (C5's 4097 points: 143 KB)
        mode |= kModeLdsQ;
        lds = lds_state_offset(nxp) + lds_q_bytes(nxp);
    } else if (!hbm && count > 1) {
        mode = lds_auto(mode, wgs, &lds);
    }
    launch_solve(hbm != nullptr, flux, dim3(wgs), dim3(threads), lds, stream, d_inst + first, count,
                 d_order_for(first, count), d_qctr, mode, max_steps, nxp, hbm);
    return hipGetLastError() == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

// The whole batch as one graph launch: reset the work counter, then one
// persistent k1_solve over all instances (largest first).  The graph is
// re-captured only when the launch shape or the step budget changes.
int Batch1D::launch_graph(int max_steps) {
    const int nxp = ((max_nx + 2 * G) + 15) / 16 * 16;
    const int wgs = workgroups > 0 ? std::min(workgroups, n) : n;
    uint8_t* hbm = nullptr;
    if (int st = mask_space(wgs, nxp, &hbm)) return st;
    if (!gexec || g_threads != threads || g_wgs != wgs || g_steps != max_steps || g_hbm != hbm || g_flux != flux) {
        if (gexec) { hipGraphExecDestroy(gexec); gexec = nullptr; }
        if (graph) { hipGraphDestroy(graph); graph = nullptr; }
        if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return AWCM_EDEVICE;
        hipLaunchKernelGGL(k1_reset_counter, dim3(1), dim3(64), 0, stream, d_qctr);
        size_t lds = hbm ? 0 : lds_bytes(nxp);
        const int mode = hbm ? 1 : lds_auto(1, wgs, &lds);
        launch_solve(hbm != nullptr, flux, dim3(wgs), dim3(threads), lds, stream, d_inst, n, d_order, d_qctr, mode,
                     max_steps, nxp, hbm);
        if (hipStreamEndCapture(stream, &graph) != hipSuccess) return AWCM_EDEVICE;
        if (hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess) return AWCM_EDEVICE;
        g_threads = threads;
        g_wgs = wgs;
        g_steps = max_steps;
        g_hbm = hbm;
        g_flux = flux;
    }
    return hipGraphLaunch(gexec, stream) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

// Device copy of every instance's state (Q, masks, levels) and of the host
// descriptors; restore() puts it back.  Used to run autotuner probes without
// changing the solve.
int Batch1D::snapshot() {
    if (snap_q.empty()) {
        snap_q.resize(n);
        snap_m.resize(n);
        for (int k = 0; k < n; k++) {
            const size_t NP = size_t(h[k].Nx) + 2 * G;
            if (dalloc(&snap_q[k], sizeof(double) * 3 * NP)) return AWCM_ENOMEM;
            if (dalloc(&snap_m[k], 5 * NP + 4 * size_t(h[k].Nx))) return AWCM_ENOMEM;
        }
    }
    for (int k = 0; k < n; k++) {
        const size_t NP = size_t(h[k].Nx) + 2 * G;
        if (hipMemcpyAsync(snap_q[k], h[k].Q, sizeof(double) * 3 * NP, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
            hipMemcpyAsync(snap_m[k], h[k].masks, 5 * NP, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
            hipMemcpyAsync(snap_m[k] + 5 * NP, h[k].levels, 4 * size_t(h[k].Nx), hipMemcpyDeviceToDevice, stream) !=
                hipSuccess)
            return AWCM_EDEVICE;
    }
    snap_h = h;
    return hipStreamSynchronize(stream) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

int Batch1D::restore() {
    if (snap_q.empty()) return AWCM_ESTATE;
    for (int k = 0; k < n; k++) {
        const size_t NP = size_t(h[k].Nx) + 2 * G;
        if (hipMemcpyAsync(h[k].Q, snap_q[k], sizeof(double) * 3 * NP, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
            hipMemcpyAsync(h[k].masks, snap_m[k], 5 * NP, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
            hipMemcpyAsync(h[k].levels, snap_m[k] + 5 * NP, 4 * size_t(h[k].Nx), hipMemcpyDeviceToDevice, stream) !=
                hipSuccess)
            return AWCM_EDEVICE;
    }
    h = snap_h;
    if (int st = push()) return st;
    return hipStreamSynchronize(stream) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

// For a launch over a sub-range (sequential mode) the order is the identity.
const int* Batch1D::d_order_for(int first, int count) {
    if (first == 0 && count == n) return d_order;
    if (!d_identity) {
        std::vector<int> id(n, 0);
        if (dalloc(&d_identity, sizeof(int) * n)) return d_order;
        hipMemcpy(d_identity, id.data(), sizeof(int) * n, hipMemcpyHostToDevice);
    }
    return d_identity;
}

// First error status over the instances.  With nonfinite_ok, an instance
// whose state broke down (status AWCM_ENONFINITE, stopped at its last finite
// time) is not an error of the call: the batch reports it per instance.
int Batch1D::check_status(bool nonfinite_ok) {
    for (int k = 0; k < n; k++) {
        const int s = h[k].status;
        if (s == 0 || (s == AWCM_ENONFINITE && nonfinite_ok)) continue;
        return s == -5 ? AWCM_ERANGE : s == AWCM_ENONFINITE ? AWCM_ENONFINITE : AWCM_EDEVICE;
    }
    return AWCM_OK;
}

}  // namespace awcm

using awcm::Batch1D;

extern "C" {

int awcm_batch_create_1d(int n, const awcm_instance_1d* params, const double* print_times, int n_print,
                         awcm_batch** out) {
    if (!out) return AWCM_EINVAL;
    *out = nullptr;
    auto* b = new (std::nothrow) awcm_batch();
    if (!b) return AWCM_ENOMEM;
    const int st = b->b1.init(n, params, print_times, n_print);
    if (st) { delete b; return st; }
    *out = b;
    return AWCM_OK;
}

void awcm_batch_destroy(awcm_batch* b) { delete b; }

int awcm_batch_upload_q(awcm_batch* b, int idx, const double* q) {
    if (!b || !q || idx < 0 || idx >= b->b1.n) return AWCM_EINVAL;
    return awcm::upload_q1(b->b1, idx, q);
}

int awcm_batch_download_q(awcm_batch* b, int idx, double* q) {
    if (!b || !q || idx < 0 || idx >= b->b1.n) return AWCM_EINVAL;
    return awcm::download_q1(b->b1, idx, q);
}

int awcm_batch_get_masks(awcm_batch* b, int idx, uint8_t* masks, int32_t* levels) {
    if (!b || idx < 0 || idx >= b->b1.n) return AWCM_EINVAL;
    return awcm::get_masks1(b->b1, idx, masks, levels);
}

int awcm_batch_set_launch(awcm_batch* b, int threads, int workgroups) {
    if (!b || threads < 64 || threads > 1024 || threads % 64 || workgroups < 0) return AWCM_EINVAL;
    b->b1.threads = threads;
    b->b1.workgroups = workgroups;
    return AWCM_OK;
}

int awcm_batch_set_flux(awcm_batch* b, int kind) {
    if (!b || (kind != AWCM_FLUX_WENOZ && kind != AWCM_FLUX_WENO5)) return AWCM_EINVAL;
    b->b1.flux = kind;
    return AWCM_OK;
}

int awcm_batch_run(awcm_batch* b, int max_steps) {
    if (!b || max_steps < 0) return AWCM_EINVAL;
    Batch1D& B = b->b1;
    int st = B.push();
    if (!st) st = B.launch_graph(max_steps);
    if (!st) st = B.pull();
    if (!st) st = B.check_status(true);
    return st;
}

int awcm_batch_run_sequential(awcm_batch* b, int max_steps) {
    if (!b || max_steps < 0) return AWCM_EINVAL;
    Batch1D& B = b->b1;
    int st = B.push();
    for (int k = 0; k < B.n && !st; k++) st = B.launch(1, max_steps, k, 1);
    if (!st) st = B.pull();
    if (!st) st = B.check_status(true);
    return st;
}

int awcm_batch_autotune(awcm_batch* b, int probe_steps, int* threads, int* workgroups, double* probe_ms) {
    if (!b || probe_steps <= 0) return AWCM_EINVAL;
    Batch1D& B = b->b1;
    // candidate launch shapes: workgroup size x number of persistent workers
    // (0 = one workgroup per instance); each is timed on the same probe
    // (probe_steps steps of every instance from the current state), which is
    // then undone, so tuning does not change the solve
    static const int kThreads[3] = {256, 512, 1024};
    static const int kWgFactor[3] = {0, 1, 2};  // 0: n workgroups; f: f x (number of CUs)
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return AWCM_EDEVICE;
    const int keep_t = B.threads, keep_w = B.workgroups;
    int st = B.push();
    if (!st) st = B.snapshot();
    if (st) return st;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return AWCM_EDEVICE;
    double best = 1e300;
    int bt = keep_t, bw = keep_w, idx = 0;
    for (int ti = 0; ti < 3 && !st; ti++) {
        for (int wi = 0; wi < 3 && !st; wi++, idx++) {
            B.threads = kThreads[ti];
            B.workgroups = kWgFactor[wi] * cus;
            if (B.workgroups >= B.n) B.workgroups = 0;
            st = B.launch_graph(probe_steps);  // capture outside the timed region
            if (!st) st = B.restore();
            if (!st && hipEventRecord(e0, B.stream) != hipSuccess) st = AWCM_EDEVICE;
            if (!st) st = B.launch_graph(probe_steps);
            if (!st && hipEventRecord(e1, B.stream) != hipSuccess) st = AWCM_EDEVICE;
            float ms = 0;
            if (!st && hipEventSynchronize(e1) != hipSuccess) st = AWCM_EDEVICE;
            if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = AWCM_EDEVICE;
            if (!st) st = B.restore();
            if (probe_ms) probe_ms[idx] = st ? -1.0 : ms;
            if (!st && ms < best) { best = ms; bt = B.threads; bw = B.workgroups; }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    B.threads = st ? keep_t : bt;
    B.workgroups = st ? keep_w : bw;
    if (threads) *threads = B.threads;
    if (workgroups) *workgroups = B.workgroups;
    return st;
}

int awcm_batch_state(awcm_batch* b, int idx, double* t, int* cnt, int* steps, long long* updates, int* status) {
    if (!b || idx < 0 || idx >= b->b1.n) return AWCM_EINVAL;
    const awcm::Inst1D& in = b->b1.h[idx];
    if (t) *t = in.t;
    if (cnt) *cnt = in.cnt;
    if (steps) *steps = in.steps;
    if (updates) *updates = in.sum_active0;
    if (status) *status = in.status;
    return AWCM_OK;
}

int awcm_batch_synchronize(awcm_batch* b) {
    if (!b) return AWCM_EINVAL;
    return hipStreamSynchronize(b->b1.stream) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

}  // extern "C"

namespace awcm {

int upload_q1(Batch1D& B, int idx, const double* q) {
    Inst1D& in = B.h[idx];
    const int NP = in.Nx + 2 * 3;
    std::vector<double> soa(size_t(3) * NP);
    for (int i = 0; i < NP; i++)
        for (int c = 0; c < 3; c++) soa[size_t(c) * NP + i] = q[size_t(i) * 3 + c];
    if (hipMemcpy(in.Q, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice) != hipSuccess) return AWCM_EDEVICE;
    return AWCM_OK;
}

int download_q1(Batch1D& B, int idx, double* q) {
    Inst1D& in = B.h[idx];
    const int NP = in.Nx + 2 * 3;
    std::vector<double> soa(size_t(3) * NP);
    if (hipStreamSynchronize(B.stream) != hipSuccess) return AWCM_EDEVICE;
    if (hipMemcpy(soa.data(), in.Q, sizeof(double) * soa.size(), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    for (int i = 0; i < NP; i++)
        for (int c = 0; c < 3; c++) q[size_t(i) * 3 + c] = soa[size_t(c) * NP + i];
    return AWCM_OK;
}

int get_masks1(Batch1D& B, int idx, uint8_t* masks, int32_t* levels) {
    Inst1D& in = B.h[idx];
    const int NP = in.Nx + 2 * 3;
    if (hipStreamSynchronize(B.stream) != hipSuccess) return AWCM_EDEVICE;
    if (masks && hipMemcpy(masks, in.masks, size_t(5) * NP, hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    if (levels) {
        std::vector<uint8_t> lv(size_t(4) * in.Nx);
        if (hipMemcpy(lv.data(), in.levels, lv.size(), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
        for (size_t k = 0; k < lv.size(); k++) levels[k] = lv[k];
    }
    return AWCM_OK;
}

// The scalar field of AWCM_1D::Prepare / WAVELET_1D::Interpolate: plane 0 of
// the instance's field, (Nx+6) doubles.
int upload_scalar1(Batch1D& B, int idx, const double* q) {
    Inst1D& in = B.h[idx];
    if (hipStreamSynchronize(B.stream) != hipSuccess) return AWCM_EDEVICE;
    return hipMemcpy(in.Q, q, sizeof(double) * (in.Nx + 6), hipMemcpyHostToDevice) == hipSuccess ? AWCM_OK
                                                                                                : AWCM_EDEVICE;
}
int download_scalar1(Batch1D& B, int idx, double* q) {
    Inst1D& in = B.h[idx];
    if (hipStreamSynchronize(B.stream) != hipSuccess) return AWCM_EDEVICE;
    return hipMemcpy(q, in.Q, sizeof(double) * (in.Nx + 6), hipMemcpyDeviceToHost) == hipSuccess ? AWCM_OK
                                                                                                : AWCM_EDEVICE;
}
// WAVELET_1D::Interpolate / InterpolateEuler on instance idx (k1_solve modes
// 4, 5, 6): the kernel reads from and to in the mask_temp and mask3 slots, so
// the instance's own masks are saved around it and put back.
int interpolate1(Batch1D& B, int idx, const uint8_t* from, const uint8_t* to, int mode) {
    Inst1D& in = B.h[idx];
    const size_t NP = size_t(in.Nx) + 6;
    std::vector<uint8_t> saved(5 * NP);
    if (hipStreamSynchronize(B.stream) != hipSuccess ||
        hipMemcpy(saved.data(), in.masks, saved.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(in.masks, from, NP, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(in.masks + 4 * NP, to, NP, hipMemcpyHostToDevice) != hipSuccess)
        return AWCM_EDEVICE;
    int st = B.push();
    if (!st) st = B.launch(mode, 0, idx, 1);
    if (!st) st = B.pull();
    if (hipStreamSynchronize(B.stream) != hipSuccess ||
        hipMemcpy(in.masks, saved.data(), saved.size(), hipMemcpyHostToDevice) != hipSuccess)
        st = st ? st : AWCM_EDEVICE;
    return st;
}

// AWCM_MANAGER_1D's statics on a caller mask (and level field) for instance
// idx (k1_solve modes 7..10): the kernel works on the mask3 and level3 slots,
// so the instance's own masks and levels are saved around it and put back.
// mode 10 (ComputeLevel) returns the level field, the others the mask.
int mask_op1(Batch1D& B, int idx, int mode, uint8_t* mask, int32_t* levels) {
    Inst1D& in = B.h[idx];
    const size_t NP = size_t(in.Nx) + 6, nl = size_t(in.Nx);
    std::vector<uint8_t> saved_m(5 * NP), saved_l(4 * nl), lb(nl, 0);
    if (levels)
        for (size_t i = 0; i < nl; i++) {
            if (levels[i] < 0 || levels[i] > in.L - 1) return AWCM_EINVAL;
            lb[i] = uint8_t(levels[i]);
        }
    if (hipStreamSynchronize(B.stream) != hipSuccess ||
        hipMemcpy(saved_m.data(), in.masks, saved_m.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(saved_l.data(), in.levels, saved_l.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(in.masks + 4 * NP, mask, NP, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(in.levels + 3 * nl, lb.data(), nl, hipMemcpyHostToDevice) != hipSuccess)
        return AWCM_EDEVICE;
    in.status = 0;
    int st = B.push();
    if (!st) st = B.launch(mode, 0, idx, 1);
    if (!st) st = B.pull();
    if (!st && in.status) st = in.status == -5 ? AWCM_ERANGE : AWCM_EDEVICE;
    if (!st) {
        if (mode == 10) {
            if (hipMemcpy(lb.data(), in.levels + 3 * nl, nl, hipMemcpyDeviceToHost) != hipSuccess) st = AWCM_EDEVICE;
            else
                for (size_t i = 0; i < nl; i++) levels[i] = lb[i];
        } else if (hipMemcpy(mask, in.masks + 4 * NP, NP, hipMemcpyDeviceToHost) != hipSuccess) {
            st = AWCM_EDEVICE;
        }
    }
    in.status = 0;
    if (hipMemcpy(in.masks, saved_m.data(), saved_m.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(in.levels, saved_l.data(), saved_l.size(), hipMemcpyHostToDevice) != hipSuccess)
        st = st ? st : AWCM_EDEVICE;
    return st;
}

// AWCM_MANAGER_1D::RemoveResidue (AWCM_MANAGER_1D.h:135-157) on an AoS field
// of ncomp components, (Nx+6) points: every component zeroed at the finest
// points outside the mask.
__global__ void k1_residue_aos(double* q, const uint8_t* mask, int Nx, int ncomp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nx || mask[i + G]) return;
    for (int c = 0; c < ncomp; c++) q[(long long)(i + G) * ncomp + c] = 0.0;
}

int remove_residue1(Batch1D& B, int idx, double* q, int ncomp, const uint8_t* mask) {
    const Inst1D& in = B.h[idx];
    const size_t NP = size_t(in.Nx) + 6;
    double* dq = nullptr;
    uint8_t* dm = nullptr;
    if (hipMalloc(&dq, sizeof(double) * NP * ncomp) != hipSuccess) return AWCM_ENOMEM;
    if (hipMalloc(&dm, NP) != hipSuccess) { hipFree(dq); return AWCM_ENOMEM; }
    int st = AWCM_OK;
    if (hipStreamSynchronize(B.stream) != hipSuccess ||
        hipMemcpy(dq, q, sizeof(double) * NP * ncomp, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dm, mask, NP, hipMemcpyHostToDevice) != hipSuccess)
        st = AWCM_EDEVICE;
    if (!st) {
        hipLaunchKernelGGL(k1_residue_aos, dim3((in.Nx + 255) / 256), dim3(256), 0, B.stream, dq, dm, in.Nx, ncomp);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(B.stream) != hipSuccess ||
            hipMemcpy(q, dq, sizeof(double) * NP * ncomp, hipMemcpyDeviceToHost) != hipSuccess)
            st = AWCM_EDEVICE;
    }
    hipFree(dq);
    hipFree(dm);
    return st;
}

int set_mask0_1(Batch1D& B, int idx, const uint8_t* m0) {
    Inst1D& in = B.h[idx];
    const int NP = in.Nx + 2 * 3;
    if (hipStreamSynchronize(B.stream) != hipSuccess) return AWCM_EDEVICE;
    return hipMemcpy(in.masks + NP, m0, NP, hipMemcpyHostToDevice) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

}  // namespace awcm
