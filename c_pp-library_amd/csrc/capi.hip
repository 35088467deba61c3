// C ABI entry points of the solver handle (include/awcm_capi.h).  A 1-D handle
// is a batch of one instance; a 2-D handle owns a Solver2D.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "awcm/physics.h"
#include "awcm_capi.h"
#include "runtime.h"

using awcm::Batch1D;
using awcm::Solver2D;

namespace awcm {

int set_schedule1(Batch1D& B, const double* pt, int npt) {
    if (npt < 0 || (npt > 0 && !pt)) return AWCM_EINVAL;
    if (hipStreamSynchronize(B.stream) != hipSuccess) return AWCM_EDEVICE;
    if (npt > B.n_print) {
        double* p = nullptr;
        if (B.dalloc(&p, sizeof(double) * npt)) return AWCM_ENOMEM;
        B.d_pt = p;
    }
    if (npt > 0 && hipMemcpy(B.d_pt, pt, sizeof(double) * npt, hipMemcpyHostToDevice) != hipSuccess) return AWCM_EDEVICE;
    B.n_print = std::max(B.n_print, npt);
    for (auto& in : B.h) { in.n_print = npt; in.print_times = B.d_pt; }
    return AWCM_OK;
}

}  // namespace awcm

namespace {

int flag_status(int flag) {
    if (flag & 2) return AWCM_ENONFINITE;
    if (flag & 1) return AWCM_ERANGE;
    return AWCM_OK;
}

// q[i] = a[i] / b[i] through the device's shared-divisor quotient
__global__ void k_divide_shared(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ q,
                                long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = awcm::Div(b[i])(a[i]);
}

}  // namespace

extern "C" {

int awcm_create_1d(int num_levels, int nx_coarse, double gamma, double dx, awcm_solver** out) {
    if (!out) return AWCM_EINVAL;
    *out = nullptr;
    auto* h = new (std::nothrow) awcm_solver();
    if (!h) return AWCM_ENOMEM;
    h->dim = 1;
    awcm_instance_1d p;
    p.num_levels = num_levels;
    p.nx_coarse = nx_coarse;
    p.gamma = gamma;
    p.dx = dx;
    p.tol = 0;
    p.cfl = 0.5;
    p.t_final = 0;
    const int st = h->b1.init(1, &p, nullptr, 0);
    if (st) { delete h; return st; }
    *out = h;
    return AWCM_OK;
}

int awcm_create_2d(int num_levels, int nx_coarse, int ny_coarse, double gamma, double dx, double x0, int bc_kind,
                   awcm_solver** out) {
    if (!out) return AWCM_EINVAL;
    *out = nullptr;
    auto* h = new (std::nothrow) awcm_solver();
    if (!h) return AWCM_ENOMEM;
    h->dim = 2;
    h->s2 = new (std::nothrow) Solver2D();
    if (!h->s2) { delete h; return AWCM_ENOMEM; }
    const int st = h->s2->init(num_levels, nx_coarse, ny_coarse, gamma, dx, x0, bc_kind);
    if (st) { delete h; return st; }
    *out = h;
    return AWCM_OK;
}

void awcm_destroy(awcm_solver* h) { delete h; }

int awcm_dims(const awcm_solver* h, int* nx, int* ny) {
    if (!h) return AWCM_EINVAL;
    if (h->dim == 1) {
        if (nx) *nx = h->b1.h[0].Nx;
        if (ny) *ny = 1;
    } else {
        if (nx) *nx = h->s2->D.Nx;
        if (ny) *ny = h->s2->D.Ny;
    }
    return AWCM_OK;
}

int awcm_upload_q(awcm_solver* h, const double* q) {
    if (!h || !q) return AWCM_EINVAL;
    return h->dim == 1 ? awcm::upload_q1(h->b1, 0, q) : h->s2->upload_q(q);
}

int awcm_download_q(awcm_solver* h, double* q) {
    if (!h || !q) return AWCM_EINVAL;
    return h->dim == 1 ? awcm::download_q1(h->b1, 0, q) : h->s2->download_q(q);
}

int awcm_set_mask0(awcm_solver* h, const uint8_t* mask0) {
    if (!h || !mask0) return AWCM_EINVAL;
    return h->dim == 1 ? awcm::set_mask0_1(h->b1, 0, mask0) : h->s2->set_mask0(mask0);
}

int awcm_get_masks(awcm_solver* h, uint8_t* masks, int32_t* levels) {
    if (!h) return AWCM_EINVAL;
    return h->dim == 1 ? awcm::get_masks1(h->b1, 0, masks, levels) : h->s2->get_masks(masks, levels);
}

int awcm_set_schedule(awcm_solver* h, const double* print_times, int n_print) {
    if (!h) return AWCM_EINVAL;
    return h->dim == 1 ? awcm::set_schedule1(h->b1, print_times, n_print) : h->s2->set_schedule(print_times, n_print);
}

int awcm_prepare_euler(awcm_solver* h, double tol) {
    if (!h) return AWCM_EINVAL;
    if (h->dim == 1) {
        Batch1D& B = h->b1;
        B.h[0].tol = tol;
        int st = B.push();
        if (!st) st = B.launch(0, 0, 0, 1);
        if (!st) st = B.pull();
        if (!st) st = B.check_status();
        return st;
    }
    Solver2D& S = *h->s2;
    S.tol_host = tol;
    int st = S.enqueue_prepare(true);
    int flag = 0;
    if (!st) st = S.get_state(nullptr, nullptr, nullptr, nullptr, &flag);
    return st ? st : flag_status(flag);
}

int awcm_set_flux(awcm_solver* h, int kind) {
    if (!h || (kind != AWCM_FLUX_WENOZ && kind != AWCM_FLUX_WENO5)) return AWCM_EINVAL;
    if (h->dim != 1) return AWCM_ESTATE;
    h->b1.flux = kind;
    return AWCM_OK;
}

int awcm_upload_scalar(awcm_solver* h, const double* q) {
    if (!h || !q) return AWCM_EINVAL;
    return h->dim == 1 ? awcm::upload_scalar1(h->b1, 0, q) : h->s2->upload_scalar(q);
}

int awcm_download_scalar(awcm_solver* h, double* q) {
    if (!h || !q) return AWCM_EINVAL;
    return h->dim == 1 ? awcm::download_scalar1(h->b1, 0, q) : h->s2->download_scalar(q);
}

int awcm_prepare(awcm_solver* h, double tol) {
    if (!h) return AWCM_EINVAL;
    if (h->dim == 1) {
        Batch1D& B = h->b1;
        B.h[0].tol = tol;
        int st = B.push();
        if (!st) st = B.launch(3, 0, 0, 1);
        if (!st) st = B.pull();
        if (!st) st = B.check_status();
        return st;
    }
    Solver2D& S = *h->s2;
    S.tol_host = tol;
    int st = S.enqueue_prepare(true, false);
    int flag = 0;
    if (!st) st = S.get_state(nullptr, nullptr, nullptr, nullptr, &flag);
    return st ? st : flag_status(flag);
}

int awcm_post_process(awcm_solver* h) {
    if (!h) return AWCM_EINVAL;
    return awcm_post_process_euler(h);
}

int awcm_interpolate(awcm_solver* h, const uint8_t* mask_from, const uint8_t* mask_to) {
    if (!h || !mask_from || !mask_to) return AWCM_EINVAL;
    if (h->dim == 1) {
        const int st = awcm::interpolate1(h->b1, 0, mask_from, mask_to, 4);
        return st ? st : h->b1.check_status();
    }
    return h->s2->interpolate(mask_from, mask_to, false);
}

int awcm_interpolate_euler(awcm_solver* h, const uint8_t* mask_from, const uint8_t* mask_to) {
    if (!h || !mask_from || !mask_to) return AWCM_EINVAL;
    if (h->dim == 1) {
        const int st = awcm::interpolate1(h->b1, 0, mask_from, mask_to, 5);
        return st ? st : h->b1.check_status();
    }
    return h->s2->interpolate(mask_from, mask_to, true);
}

int awcm_interpolate_all_grid(awcm_solver* h, const uint8_t* mask_from) {
    if (!h || !mask_from) return AWCM_EINVAL;
    if (h->dim == 1) {
        std::vector<uint8_t> all(size_t(h->b1.h[0].Nx) + 6, 1);
        const int st = awcm::interpolate1(h->b1, 0, mask_from, all.data(), 6);
        return st ? st : h->b1.check_status();
    }
    return h->s2->interpolate_all(mask_from);
}

int awcm_run_scalar(awcm_solver* h, double tol, double dt, int n_steps) {
    if (!h || n_steps < 0 || !(dt >= 0)) return AWCM_EINVAL;
    if (h->dim != 2) return AWCM_ESTATE;
    return h->s2->advect(tol, dt, n_steps);
}

int awcm_post_process_euler(awcm_solver* h) {
    if (!h) return AWCM_EINVAL;
    if (h->dim == 1) {
        Batch1D& B = h->b1;
        int st = B.push();
        if (!st) st = B.launch(2, 0, 0, 1);
        if (!st) st = B.pull();
        return st;
    }
    int st = h->s2->enqueue_post();
    if (!st) st = hipStreamSynchronize(h->s2->stream) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
    return st;
}

// AWCM_MANAGER_1D / AWCM_MANAGER_2D statics on caller arrays
int awcm_add_adjacent_zone(awcm_solver* h, uint8_t* mask) {
    if (!h || !mask) return AWCM_EINVAL;
    return h->dim == 1 ? awcm::mask_op1(h->b1, 0, 7, mask, nullptr)
                       : h->s2->mask_op(Solver2D::MASK_OP_ADJACENT, mask, nullptr);
}

int awcm_check_reconstruction(awcm_solver* h, uint8_t* mask) {
    if (!h || !mask) return AWCM_EINVAL;
    return h->dim == 1 ? awcm::mask_op1(h->b1, 0, 8, mask, nullptr)
                       : h->s2->mask_op(Solver2D::MASK_OP_CHECK, mask, nullptr);
}

int awcm_add_point_for_differentiation(awcm_solver* h, uint8_t* mask, const int32_t* level) {
    if (!h || !mask || !level) return AWCM_EINVAL;
    int32_t* lv = const_cast<int32_t*>(level);  // read only: the ADD_POINT path does not write it
    return h->dim == 1 ? awcm::mask_op1(h->b1, 0, 9, mask, lv)
                       : h->s2->mask_op(Solver2D::MASK_OP_ADD_POINT, mask, lv);
}

int awcm_compute_level(awcm_solver* h, int32_t* level, const uint8_t* mask) {
    if (!h || !mask || !level) return AWCM_EINVAL;
    uint8_t* m = const_cast<uint8_t*>(mask);  // read only: the LEVEL path does not write it
    return h->dim == 1 ? awcm::mask_op1(h->b1, 0, 10, m, level) : h->s2->mask_op(Solver2D::MASK_OP_LEVEL, m, level);
}

int awcm_remove_residue(awcm_solver* h, double* q, int ncomp, const uint8_t* mask) {
    if (!h || !q || !mask) return AWCM_EINVAL;
    if (ncomp != 1 && ncomp != (h->dim == 1 ? 3 : 4)) return AWCM_EINVAL;
    return h->dim == 1 ? awcm::remove_residue1(h->b1, 0, q, ncomp, mask) : h->s2->remove_residue(q, ncomp, mask);
}

int awcm_run(awcm_solver* h, double tol, double cfl, double t_final, int max_steps, double* t, int* cnt,
             int* steps_done, long long* updates) {
    // cnt indexes the print schedule on the device: a negative one is refused
    if (!h || !t || !cnt || *cnt < 0 || max_steps < 0 || !(cfl > 0)) return AWCM_EINVAL;
    if (h->dim == 1) {
        Batch1D& B = h->b1;
        awcm::Inst1D& in = B.h[0];
        in.tol = tol; in.cfl = cfl; in.t_final = t_final; in.t = *t; in.cnt = *cnt;
        in.steps = 0; in.sum_active0 = 0; in.status = 0;
        int st = B.push();
        if (!st) st = B.launch(1, max_steps, 0, 1);
        if (!st) st = B.pull();
        if (!st) {
            *t = in.t; *cnt = in.cnt;
            if (steps_done) *steps_done = in.steps;
            if (updates) *updates = in.sum_active0;
            st = B.check_status();
        }
        return st;
    }
    Solver2D& S = *h->s2;
    S.tol_host = tol;
    int st = S.set_state(*t, *cnt, t_final, tol, cfl);
    // Queue the steps in chunks; once t >= t_final the queued steps are no-ops.
    const int chunk = 32;
    int done = 0;
    while (!st && done < max_steps) {
        const int k = std::min(chunk, max_steps - done);
        for (int s = 0; s < k && !st; s++) st = S.launch_step();
        done += k;
        double tt = 0;
        int flag = 0;
        if (!st) st = S.get_state(&tt, nullptr, nullptr, nullptr, &flag);
        if (!st) st = flag_status(flag);
        if (!st && !(tt < t_final)) break;
    }
    if (!st) {
        int nsteps = 0;
        long long upd = 0;
        st = S.get_state(t, cnt, &nsteps, &upd, nullptr);
        if (steps_done) *steps_done = nsteps;
        if (updates) *updates = upd;
    }
    return st;
}

int awcm_run_timed(awcm_solver* h, double tol, double cfl, double t_final, int n_steps, double* t, int* cnt,
                   int* steps_done, long long* updates, double* elapsed_ms) {
    if (!h || !t || !cnt || *cnt < 0 || n_steps < 0 || !(cfl > 0)) return AWCM_EINVAL;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return AWCM_EDEVICE;
    int st = AWCM_OK;
    float ms = 0;
    if (h->dim == 1) {
        Batch1D& B = h->b1;
        awcm::Inst1D& in = B.h[0];
        in.tol = tol; in.cfl = cfl; in.t_final = t_final; in.t = *t; in.cnt = *cnt;
        in.steps = 0; in.sum_active0 = 0; in.status = 0;
        st = B.push();
        if (!st && hipStreamSynchronize(B.stream) != hipSuccess) st = AWCM_EDEVICE;
        if (!st && hipEventRecord(e0, B.stream) != hipSuccess) st = AWCM_EDEVICE;
        if (!st) st = B.launch(1, n_steps, 0, 1);
        if (!st && hipEventRecord(e1, B.stream) != hipSuccess) st = AWCM_EDEVICE;
        if (!st) st = B.pull();
        if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = AWCM_EDEVICE;
        if (!st) {
            *t = in.t; *cnt = in.cnt;
            if (steps_done) *steps_done = in.steps;
            if (updates) *updates = in.sum_active0;
            st = B.check_status();
        }
    } else {
        Solver2D& S = *h->s2;
        S.tol_host = tol;
        st = S.set_state(*t, *cnt, t_final, tol, cfl);
        if (!st && hipEventRecord(e0, S.stream) != hipSuccess) st = AWCM_EDEVICE;
        for (int s = 0; s < n_steps && !st; s++) st = S.launch_step();
        if (!st && hipEventRecord(e1, S.stream) != hipSuccess) st = AWCM_EDEVICE;
        int nsteps = 0, flag = 0;
        long long upd = 0;
        if (!st) st = S.get_state(t, cnt, &nsteps, &upd, &flag);
        if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = AWCM_EDEVICE;
        if (!st) {
            if (steps_done) *steps_done = nsteps;
            if (updates) *updates = upd;
            st = flag_status(flag);
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (elapsed_ms) *elapsed_ms = ms;
    return st;
}

int awcm_profile_enable(awcm_solver* h, int enable) {
    if (!h) return AWCM_EINVAL;
    if (h->dim != 2) return AWCM_ESTATE;
    Solver2D& S = *h->s2;
    if (hipStreamSynchronize(S.stream) != hipSuccess) return AWCM_EDEVICE;
    S.prof = enable != 0;
    S.prof_clear();
    awcm::Scal2 sc;
    if (hipMemcpy(&sc, S.sc, sizeof(sc), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    sc.sumW = 0;
    sc.sumC = 0;
    sc.sumF = 0;
    return hipMemcpy(S.sc, &sc, sizeof(sc), hipMemcpyHostToDevice) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

int awcm_profile_read(awcm_solver* h, int kernel, double* total_ms, long long* launches, long long* points) {
    if (!h || kernel < 0 || kernel > 2) return AWCM_EINVAL;
    if (h->dim != 2) return AWCM_ESTATE;
    Solver2D& S = *h->s2;
    if (hipStreamSynchronize(S.stream) != hipSuccess) return AWCM_EDEVICE;
    double tot = 0;
    for (auto& pr : S.prof_ev[kernel]) {
        float ms = 0;
        if (!pr.first || !pr.second || hipEventElapsedTime(&ms, pr.first, pr.second) != hipSuccess) return AWCM_EDEVICE;
        tot += ms;
    }
    awcm::Scal2 sc;
    if (hipMemcpy(&sc, S.sc, sizeof(sc), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    if (total_ms) *total_ms = tot;
    if (launches) *launches = (long long)S.prof_ev[kernel].size();
    if (points) *points = kernel == 0 ? sc.sumF : kernel == 1 ? sc.sumW + sc.sumC : 0;
    return AWCM_OK;
}

int awcm_set_tuning(awcm_solver* h, int param, int value) {
    if (!h) return AWCM_EINVAL;
    if (h->dim != 2) return AWCM_ESTATE;
    Solver2D& S = *h->s2;
    if (hipStreamSynchronize(S.stream) != hipSuccess) return AWCM_EDEVICE;
    if (param == AWCM_TUNE_TAIL_CELLS) {
        if (value < 0) return AWCM_EINVAL;
        return S.set_tail(value);
    }
    if (param == AWCM_TUNE_GRAPH) {
        S.use_graph = value != 0;
        return AWCM_OK;
    }
    return AWCM_EINVAL;
}

int awcm_divide_shared(const double* a, const double* b, double* q, long long n) {
    if (n < 0 || (n > 0 && (!a || !b || !q))) return AWCM_EINVAL;
    if (n == 0) return AWCM_OK;
    int dev = 0;  // the caller's current device, restored on return
    if (hipGetDevice(&dev) != hipSuccess || hipSetDevice(0) != hipSuccess) return AWCM_EDEVICE;
    double* d = nullptr;
    const size_t bytes = sizeof(double) * size_t(n);
    int st = hipMalloc(&d, 3 * bytes) == hipSuccess ? AWCM_OK : AWCM_ENOMEM;
    if (st == AWCM_OK && (hipMemcpy(d, a, bytes, hipMemcpyHostToDevice) != hipSuccess ||
                          hipMemcpy(d + n, b, bytes, hipMemcpyHostToDevice) != hipSuccess))
        st = AWCM_EDEVICE;
    if (st == AWCM_OK) {
        hipLaunchKernelGGL(k_divide_shared, dim3(unsigned((n + 255) / 256)), dim3(256), 0, 0, d, d + n, d + 2 * n, n);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(0) != hipSuccess ||
            hipMemcpy(q, d + 2 * n, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            st = AWCM_EDEVICE;
    }
    if (d && hipFree(d) != hipSuccess && st == AWCM_OK) st = AWCM_EDEVICE;
    if (hipSetDevice(dev) != hipSuccess && st == AWCM_OK) st = AWCM_EDEVICE;
    return st;
}

int awcm_synchronize(awcm_solver* h) {
    if (!h) return AWCM_EINVAL;
    hipStream_t s = h->dim == 1 ? h->b1.stream : h->s2->stream;
    return hipStreamSynchronize(s) == hipSuccess ? AWCM_OK : AWCM_EDEVICE;
}

}  // extern "C"
