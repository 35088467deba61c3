// CPU emulation of the 1-D HIP kernels (k1_solve with tape recording, and
// k1_adjoint), built with g++ against tests/native/emu/hip/hip_runtime.h (one
// emulated thread per workgroup) or against tests/native/emu_mt/hip/
// hip_runtime.h with EMU_THREADS threads (a workgroup of real threads and
// barriers, for ThreadSanitizer).  The device code is compiled unchanged;
// only the launch is replaced by emu_launch.  Test infrastructure only: it
// lets the CPU suite check the adjoint's device code against finite
// differences, and its barrier discipline, without a GPU.
#include <vector>

#ifndef EMU_THREADS
#define EMU_THREADS 1
#endif

// the reverse sweep's grid launches become loops over the work items, one
// emulated thread per workgroup of one item
#define AWCM_REV_LAUNCH(kern, items, ...)                                       \
    do {                                                                         \
        blockDim.x = 1;                                                          \
        threadIdx.x = 0;                                                         \
        for (blockIdx.x = 0; blockIdx.x < unsigned(items); blockIdx.x++) kern(__VA_ARGS__); \
        blockIdx.x = 0;                                                          \
    } while (0)

// the one-workgroup reverse launches run on the emulated workgroup
#define AWCM_REV_LAUNCH_WG(kern, ...) emu_launch(EMU_THREADS, [&] { kern(__VA_ARGS__); })
#include "awcm1d.hip"
#include "adjoint1d_rev.h"

namespace awcm {
uint8_t sm[1 << 21];         // dynamic LDS of the emulated workgroup (k1_adjoint)
__attribute__((aligned(16))) uint8_t lds_masks[1 << 21];  // dynamic LDS of the emulated workgroup (k1_solve<false>)
}

using namespace awcm;

namespace {

struct Emu {
    Inst1D in{};
    std::vector<double> Q, Qc, Qt, F, pt;
    std::vector<uint8_t> masks, levels, tape_m;
    std::vector<double> tape_s, ckpt, seg, Qb, Xb, Cb, Gb, FV;
    int steps = 0, K = 1;
    int step = 0;  // the step being reversed
    int NP() const { return in.Nx + 6; }
    int nxp() const { return (NP() + 15) / 16 * 16; }
};

// Poison: buffers the device side leaves uninitialised (hipMalloc) are filled
// with a NaN pattern, and the emulated LDS with 0xA5 bytes before every
// launch, so a read of anything not written first shows up in the results.
bool poison = false;
double nan_pattern() {
    const uint64_t b = 0x7ff4dead7ff4deadULL;
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}
void poison_lds() {
    if (!poison) return;
    std::memset(sm, 0xA5, sizeof(sm));
    std::memset(lds_masks, 0xA5, sizeof(lds_masks));
}
template <class T>
void fill_new(std::vector<T>& v, size_t n, T zero) {
    if (poison) {
        if constexpr (sizeof(T) == 8) v.assign(n, nan_pattern());
        else v.assign(n, T(0xA5));
    } else {
        v.assign(n, zero);
    }
}

// with masks_in_hbm, k1_solve<true> keeps the masks in a heap slice instead
// of the emulated LDS (the path of instances too long for one workgroup's LDS)
bool masks_in_hbm = false;
// with lds_auto, the launch is a batched one of one workgroup per CU: the whole
// budget, and the kernel places each instance's state itself (kModeLdsAuto)
bool lds_auto = false;
int stage_flux = kRecWenoZ;  // face fluxes of the stages (kRecWenoZ / kRecWeno5)

// a workgroup's LDS beside the kernels' static LDS (a few hundred bytes; the
// device runtime asks for the exact figure): the single-instance launches put
// the whole state, or Q alone, in LDS when it fits, as Batch1D::launch does
constexpr size_t kEmuLdsCap = 163840 - 512;
int lds_mode(int nxp) {
    if (lds_state_offset(nxp) + lds_state_bytes(nxp) <= kEmuLdsCap) return kModeLdsState;
    if (lds_state_offset(nxp) + lds_q_bytes(nxp) <= kEmuLdsCap) return kModeLdsQ;
    return 0;
}

template <int FLUX>
void launch_solve_f(Emu& e, int mode, int max_steps) {
    int order = 0, qctr = 0;
    poison_lds();
    if (!masks_in_hbm) mode |= lds_auto ? kModeLdsAuto | int((kEmuLdsCap / 16) << 16) : lds_mode(e.nxp());
    if (masks_in_hbm) {
        std::vector<uint8_t> hbm(size_t(11) * e.nxp(), 0xA5);
        emu_launch(EMU_THREADS,
                   [&] { k1_solve<true, FLUX, false>(&e.in, 1, &order, &qctr, mode, max_steps, e.nxp(), hbm.data()); });
    } else {
        emu_launch(EMU_THREADS,
                   [&] { k1_solve<false, FLUX, false>(&e.in, 1, &order, &qctr, mode, max_steps, e.nxp(), nullptr); });
    }
}
void launch_solve(Emu& e, int mode, int max_steps) {
    if (stage_flux == kRecWeno5) launch_solve_f<kRecWeno5>(e, mode, max_steps);
    else launch_solve_f<kRecWenoZ>(e, mode, max_steps);
}

}  // namespace

extern "C" {

void emu_set_masks_in_hbm(int on) { masks_in_hbm = on != 0; }
void emu_set_lds_auto(int on) { lds_auto = on != 0; }
void emu_set_flux(int kind) { stage_flux = kind == 1 ? kRecWeno5 : kRecWenoZ; }

void* emu_create(int L, int nxc, double gamma, double dx, const double* pt, int npt) {
    auto* e = new Emu();
    Inst1D& in = e->in;
    in.L = L;
    in.nxc = nxc;
    in.Nx = (nxc - 1) * (1 << (L - 1)) + 1;
    in.gamma = gamma;
    in.dx = dx;
    in.cnt = 1;
    const int NP = e->NP();
    e->Q.assign(3 * NP, 0.0);
    e->Qc = e->Q;
    e->Qt = e->Q;
    fill_new(e->F, 3 * size_t(in.Nx + 1), 0.0);  // not cleared by the device runtime either
    e->masks.assign(5 * NP, 0);
    std::fill(e->masks.begin() + NP, e->masks.begin() + 2 * NP, 1);
    e->levels.assign(4 * in.Nx, 0);
    e->pt.assign(pt, pt + npt);
    in.Q = e->Q.data(); in.Qc = e->Qc.data(); in.Qt = e->Qt.data(); in.F = e->F.data();
    in.masks = e->masks.data(); in.levels = e->levels.data();
    in.print_times = e->pt.data(); in.n_print = npt;
    return e;
}

void emu_destroy(void* h) { delete static_cast<Emu*>(h); }

void emu_set_q(void* h, const double* q) {
    auto* e = static_cast<Emu*>(h);
    const int NP = e->NP();
    for (int i = 0; i < NP; i++)
        for (int c = 0; c < 3; c++) e->Q[c * NP + i] = q[i * 3 + c];
}

void emu_get_q(void* h, double* q) {
    auto* e = static_cast<Emu*>(h);
    const int NP = e->NP();
    for (int i = 0; i < NP; i++)
        for (int c = 0; c < 3; c++) q[i * 3 + c] = e->Q[c * NP + i];
}

// A recording from the instance's current state, masks and levels, as
// awcm_adjoint_record does: t and the print counter are the caller's.
int emu_record_from(void* h, double tol, double cfl, double t_final, int max_steps, int K, double t0, int cnt0) {
    auto* e = static_cast<Emu*>(h);
    Inst1D& in = e->in;
    in.t = t0; in.cnt = cnt0; in.steps = 0; in.sum_active0 = 0; in.status = 0;
    const long long NP = e->NP();
    const long long stride = adj_tape_stride(in.Nx);
    fill_new(e->tape_m, size_t(stride) * max_steps, uint8_t(0));
    fill_new(e->tape_s, 4 * size_t(max_steps), 0.0);
    fill_new(e->ckpt, 3 * NP * size_t(max_steps / K + 1), 0.0);
    e->K = K;
    in.tol = tol; in.cfl = cfl; in.t_final = t_final;
    in.tape_m = e->tape_m.data(); in.tape_s = e->tape_s.data(); in.ckpt = e->ckpt.data();
    in.tape_stride = stride; in.tape_n = 0; in.tape_cap = max_steps; in.ckpt_every = K;
    launch_solve(*e, 1, max_steps);
    e->steps = in.tape_n;
    in.tape_m = nullptr; in.tape_s = nullptr; in.ckpt = nullptr;
    in.tape_n = 0; in.tape_cap = 0;
    return e->steps;
}

// Each recording of a fresh run: t = 0, print counter 1, the masks and levels
// of a new solver (mask0 all ones, AWCM_1D.h:23-34).
int emu_record(void* h, double tol, double cfl, double t_final, int max_steps, int K) {
    auto* e = static_cast<Emu*>(h);
    const int NPr = e->NP();
    std::fill(e->masks.begin(), e->masks.end(), 0);
    std::fill(e->masks.begin() + NPr, e->masks.begin() + 2 * NPr, 1);
    std::fill(e->levels.begin(), e->levels.end(), 0);
    return emu_record_from(h, tol, cfl, t_final, max_steps, K, 0.0, 1);
}

void emu_set_poison(int on) { poison = on != 0; }

static Adj1DArgs args(Emu& e, int seg_steps) {
    Adj1DArgs A{};
    const long long NP = e.NP();
    A.tape_m = e.tape_m.data();
    A.tape_s = e.tape_s.data();
    A.ckpt = e.ckpt.data();
    A.tape_stride = adj_tape_stride(e.in.Nx);
    A.n_steps = e.steps;
    A.ckpt_every = e.K;
    A.seg_steps = seg_steps;
    fill_new(e.seg, 9 * NP * size_t(seg_steps), 0.0);
    fill_new(e.Qb, 3 * NP, 0.0);
    fill_new(e.Xb, 3 * NP, 0.0);
    fill_new(e.Cb, 3 * NP, 0.0);
    fill_new(e.Gb, 3 * NP, 0.0);
    fill_new(e.FV, 18 * size_t(e.in.Nx + 1), 0.0);
    A.seg = e.seg.data();
    A.step = &e.step;
    return A;
}

void emu_replay(void* h) {
    auto* e = static_cast<Emu*>(h);
    const Adj1DArgs A = args(*e, e->K);
    poison_lds();
    const int lq = lds_state_offset(e->nxp()) + lds_q_bytes(e->nxp()) <= kEmuLdsCap ? 1 : 0;
    emu_launch(EMU_THREADS, [&] { k1_adjoint(&e->in, A, 0, e->nxp(), lq); });
}

// awcm_adjoint's sequence: for every step from the last, the segment
// recompute (one workgroup, when the step ends its segment), the reverse
// sweep's grid kernels, and the step counter moved back.
void emu_adjoint(void* h, const double* gf, int seg_steps, double* g0) {
    auto* e = static_cast<Emu*>(h);
    const int NP = e->NP();
    std::vector<double> Qsave = e->Q;
    const int K = seg_steps ? seg_steps : e->K;
    const Adj1DArgs A = args(*e, K);
    // the C ABI uploads the whole adjoint state, ghosts as zeros
    std::fill(e->Qb.begin(), e->Qb.end(), 0.0);
    for (int i = 3; i < NP - 3; i++)
        for (int c = 0; c < 3; c++) e->Qb[c * NP + i] = gf[i * 3 + c];
    AdjRev R{};
    R.Nx = e->in.Nx; R.L = e->in.L; R.NP = NP; R.K = K;
    R.gamma = e->in.gamma; R.dx = e->in.dx;
    R.tape_m = e->tape_m.data(); R.stride = adj_tape_stride(e->in.Nx); R.tape_s = e->tape_s.data();
    R.seg = e->seg.data();
    R.Qb = e->Qb.data(); R.Xb = e->Xb.data(); R.Cb = e->Cb.data(); R.Gb = e->Gb.data(); R.FV = e->FV.data();
    R.step = &e->step;
    const int lq = lds_state_offset(e->nxp()) + lds_q_bytes(e->nxp()) <= kEmuLdsCap ? 1 : 0;
    for (e->step = e->steps - 1; e->step >= 0;) {
        poison_lds();
        emu_launch(EMU_THREADS, [&] { k1_adjoint(&e->in, A, 2, e->nxp(), lq); });
        launch_reverse_step(R, nullptr);
        AWCM_REV_LAUNCH(k1r_next, 1, &e->step);
    }
    e->Q = Qsave;
    e->in.Q = e->Q.data();
    for (int i = 0; i < NP; i++)
        for (int c = 0; c < 3; c++) g0[i * 3 + c] = (i < 3 || i >= NP - 3) ? 0.0 : e->Qb[c * NP + i];
}

// The literal lifting weights of the reverse sweep equal wavelet_weights'.
int emu_check_weights() {
    double W[4][5];
    wavelet_weights(W);
    for (int cc = 0; cc < 4; cc++)
        for (int kc = 1; kc <= 3; kc++)
            if (W[cc][kc] != wlift(cc, kc)) return 0;
    return 1;
}

int emu_status(void* h) { return static_cast<Emu*>(h)->in.status; }

}  // extern "C"
