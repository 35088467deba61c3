// Host object of one 2-D AWCM solve: device buffers, geometry and the kernel
// sequence of one step (see awcm2d.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace awcm {

// An integer lattice of the MRA grid (GRID_INT_2D, MRA_GRID_UNIFORM_2D.h:10-60).
struct Lat {
    int is, js, ie, je, dx, dy, ni, nj;
};

struct Dims2 {
    int Nx, Ny, P, L, maxlev;
    long long plane;  // doubles per state plane, P * (Ny + 6)
    int PR;           // row pitch of the reference layout, Nx + 6 (host side)
};

struct Wts {
    double w[4][5];  // wavelet weight(i, j) at [i][j + 1]
};

// Device scalars of the time loop (one allocation).
struct Scal2 {
    unsigned long long vmax;  // bits of max(SRx + SRy) over mask3
    double t, dt, t_final, tol, cfl;
    long long sum_active0;
    int cnt, n_print, steps, skip, do_print, flag;
    int dirty;                // Q may be non-zero outside mask0: next residue pass is full
    int residue_full;         // this step's residue pass zeroes every inactive point
    int cntW[3], cntC[3];     // stage work-list lengths (mask0, mask1, mask2): finest / coarser points
    int cntXS[3], cntYS[3];   // finest points whose left / bottom neighbour is not a finest point
    long long sumW, sumC, sumF;  // stage points and faces processed, summed over steps
};

struct Solver2D {
    Dims2 D{};
    int nxc = 0, nyc = 0, bc_kind = 0;
    double gamma = 1.4, dx = 0, x0 = 0;
    double tol_host = 0;  // wavelet threshold of the queued steps
    Lat sub[16][3]{};
    Lat grd[16]{};
    Wts W{};
    hipStream_t stream = nullptr;
    // Side stream of PrepareEuler, forked from and joined back into `stream` by
    // events (graph capture records it as a parallel branch): the interpolation
    // of momentum and energy (planes 1..3) runs beside the inverse transform of
    // density (plane 0); both only read mask3 and mask_temp.
    hipStream_t s_int = nullptr;
    hipEvent_t ev_m3 = nullptr, ev_int = nullptr;
    std::vector<void*> allocs;

    double *Q = nullptr, *Q1 = nullptr, *Q2 = nullptr;  // 4 planes each (stage 3 writes Q directly)
    uint8_t *mt = nullptr, *m[4] = {}, *snap = nullptr;
    uint8_t* tmp = nullptr;      // mask_from of interpolate_all / interpolate
    uint8_t* work = nullptr;     // work mask of interpolate
    uint8_t* lv[4] = {};
    double* alpha = nullptr;     // 3 sets of (Ny alpha_x, Nx alpha_y)
    double* apart = nullptr;     // alpha partials per tile (row, column and CFL maxima)
    uint32_t* lists = nullptr;   // 12 lists of Nx*Ny packed (j << 16 | i): W, C, XS, YS per stage mask
    double *FX = nullptr, *FY = nullptr;  // face fluxes: right / top face of each point, 4 planes
    double* pt = nullptr;        // print times
    Scal2* sc = nullptr;
    int grid_blocks = 0;         // persistent grid for list kernels
    int weno_threads = 128;

    // whole-step HIP graph (captured once per threshold value, then replayed)
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double graph_tol = -1;
    int launch_step();  // one step: graph replay, or eager when profiling

    // per-kernel HIP-event profiling
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev[3];
    hipEvent_t mark();
    void prof_clear();

    ~Solver2D();
    int init(int L, int nxc, int nyc, double gamma, double dx, double x0, int bc_kind);
    int dalloc(void* pp, size_t bytes);
    // type 0: finest-level points, 1: coarser points, 2: x-run starts, 3: y-run starts
    uint32_t* list(int k, int type) const { return lists + size_t(4 * k + type) * size_t(D.Nx) * D.Ny; }
    double* ax(int s) const { return alpha + size_t(s) * (D.Nx + D.Ny); }
    double* ay(int s) const { return alpha + size_t(s) * (D.Nx + D.Ny) + D.Ny; }

    int mark_dirty();
    int upload_q(const double* q);
    int download_q(double* q);
    int set_mask0(const uint8_t* m0);
    int get_masks(uint8_t* masks, int32_t* levels);
    int set_schedule(const double* p, int n);
    int set_state(double t, int cnt, double t_final, double tol, double cfl);
    int get_state(double* t, int* cnt, int* steps, long long* updates, int* flag);
    int interpolate_all(const uint8_t* mask_from);  // synchronous
    int interpolate(const uint8_t* mask_from, const uint8_t* mask_to, bool euler);  // synchronous
    int upload_scalar(const double* q);
    int download_scalar(double* q);
    // AWCM_MANAGER_2D statics on caller arrays (synchronous; see mask_op)
    enum { MASK_OP_ADJACENT = 0, MASK_OP_CHECK = 1, MASK_OP_ADD_POINT = 2, MASK_OP_LEVEL = 3 };
    int mask_op(int op, uint8_t* mask, int32_t* levels);
    int remove_residue(double* q, int ncomp, const uint8_t* mask);
    // the scalar box-advection driver's loop (synchronous); its stage planes
    // A2, A1, A0 are allocated, zeroed, on first use and kept
    double* adv = nullptr;
    int advect(double tol, double dt, int n_steps);

    // kernel sequences (enqueued on `stream`, no host synchronisation)
    int tail = -1;            // levels 0..tail of each cascade run in one workgroup
    int set_tail(int tail_cells);
    int enqueue_begin(bool force, bool copy_mt, bool prepare_only, bool clear_masks);
    int enqueue_prepare(bool force, bool euler = true);
    int enqueue_step();       // begin + prepare + dt + 3 stages + combine
    int enqueue_post();
};

}  // namespace awcm
