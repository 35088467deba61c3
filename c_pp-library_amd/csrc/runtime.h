// Host-side runtime objects behind the C ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "awcm_capi.h"
#include "awcm_internal.h"
#include "solver2d.h"

namespace awcm {

// N independent 1-D solves, one workgroup each (one instance = plain solver).
struct Batch1D {
    int n = 0;
    int max_nx = 0;
    int n_print = 0;
    int threads = 256;
    int flux = 0;                 // stage face fluxes of every instance: AWCM_FLUX_WENOZ / AWCM_FLUX_WENO5
    int workgroups = 0;
    std::vector<Inst1D> h;        // host mirror of the descriptors
    Inst1D* d_inst = nullptr;
    int* d_order = nullptr;
    int* d_identity = nullptr;
    int* d_qctr = nullptr;
    double* d_pt = nullptr;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    // graph of the whole-batch launch (counter reset + solve kernel), rebuilt
    // when its launch shape or step budget changes
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_threads = 0, g_wgs = 0, g_steps = -1, g_flux = 0;
    const uint8_t* g_hbm = nullptr;
    // masks and levels of instances too long for one workgroup's LDS: an
    // 11 * nxp-byte slice per workgroup in HBM (grown on demand, never shrunk)
    size_t lds_cap = 0;           // dynamic LDS one k1_solve workgroup may use
    bool lds_state = true;        // single-instance launches keep the state in LDS when it fits
    int ncu = 0;                  // compute units of the device
    uint8_t* d_hbm_masks = nullptr;
    size_t hbm_masks_bytes = 0;
    // device snapshot of every instance's state (autotuner probes)
    std::vector<double*> snap_q;
    std::vector<uint8_t*> snap_m;
    std::vector<Inst1D> snap_h;

    ~Batch1D();
    int init(int n, const awcm_instance_1d* params, const double* pt, int npt);
    int dalloc(void* pp, size_t bytes);
    int push();
    int pull();
    int launch(int mode, int max_steps, int first, int count);
    int launch_graph(int max_steps);
    int lds_auto(int mode, int wgs, size_t* lds) const;
    const int* d_order_for(int first, int count);
    int mask_space(int wgs, int nxp, uint8_t** hbm);
    int check_status(bool nonfinite_ok = false);
    int snapshot();
    int restore();
};

int upload_q1(Batch1D& B, int idx, const double* q);
int download_q1(Batch1D& B, int idx, double* q);
int get_masks1(Batch1D& B, int idx, uint8_t* masks, int32_t* levels);
int upload_scalar1(Batch1D& B, int idx, const double* q);
int download_scalar1(Batch1D& B, int idx, double* q);
int interpolate1(Batch1D& B, int idx, const uint8_t* from, const uint8_t* to, int mode);
int set_mask0_1(Batch1D& B, int idx, const uint8_t* m0);
int set_schedule1(Batch1D& B, const double* pt, int npt);
int mask_op1(Batch1D& B, int idx, int mode, uint8_t* mask, int32_t* levels);
int remove_residue1(Batch1D& B, int idx, double* q, int ncomp, const uint8_t* mask);

// Tape and work buffers of the 1-D discrete adjoint (one per handle).
struct Adj1DHost {
    std::vector<void*> allocs;
    uint8_t* tape_m = nullptr;
    double* tape_s = nullptr;
    double* ckpt = nullptr;
    long long stride = 0;
    int cap = 0, steps = 0, K = 0;
    double* seg = nullptr;
    int seg_cap = 0;      // steps the segment buffer holds
    int seg_steps = 0;    // steps per recompute segment of the current sweep
    double *Qb = nullptr, *Xb = nullptr, *Cb = nullptr, *Gb = nullptr, *FV = nullptr, *Qsave = nullptr;
    int* d_step = nullptr;  // the step being reversed (device)
    // graph of one reversed step (segment recompute, reverse kernels, step
    // counter), for segment length g_K over segment buffer g_seg
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_K = 0;
    const double* g_seg = nullptr;
    ~Adj1DHost();
    int dalloc(void* pp, size_t bytes);
};

}  // namespace awcm

struct awcm_batch {
    awcm::Batch1D b1;
};

struct awcm_solver {
    int dim = 1;
    awcm::Batch1D b1;       // dim 1
    awcm::Solver2D* s2 = nullptr;  // dim 2
    awcm::Adj1DHost* adj = nullptr;  // dim 1, after awcm_adjoint_record
    ~awcm_solver() { delete s2; delete adj; }
};
