// awcm_run_1d_noalloc: the 1-D time loop of awcm_run for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops).  It
// runs the same kernel (k1_solve, one workgroup, masks in LDS) on caller-owned
// device buffers for the state, masks and levels; everything else the solve
// needs -- the instance descriptor, the work-queue order and counter, the
// output-time schedule, the RK copy Qc, the stage input Qt and the face
// fluxes F -- lives in one file-scope device array, `scratch`, of a fixed size.
//
// Each GPU holds its own copy of `scratch` for the life of the process.  Host
// code never passes the array's name to a kernel (on the host it is a stand-in,
// not a device address): the device address is taken once per device with
// hipGetSymbolAddress and used everywhere instead.  Calls on one GPU share the
// array, so calls that overlap on one GPU race; calls run on the null stream
// and return after the solve, so calls from one host thread never overlap.

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "awcm_capi.h"
#include "awcm_internal.h"

namespace awcm {

template <bool kHbm, int FLUX, bool SCALAR>
__global__ void k1_solve(Inst1D* insts, int n_inst, const int* order, int* qctr, int mode, int max_steps, int nxp,
                         uint8_t* hbm);

namespace {

constexpr int G = 3;
constexpr int kMaxNx = AWCM_NOALLOC_MAX_NX;       // finest points per solve
constexpr int kMaxPrint = AWCM_NOALLOC_MAX_PRINT;  // output times per call
constexpr size_t kMaxNP = size_t(kMaxNx) + 2 * G;

// Byte offsets inside `scratch` (each region 256-byte aligned).
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kOffInst = 0;
constexpr size_t kOffOrder = kOffInst + align256(sizeof(Inst1D));  // int order[1], int counter
constexpr size_t kOffPrint = kOffOrder + 256;
constexpr size_t kOffQc = kOffPrint + align256(sizeof(double) * kMaxPrint);
constexpr size_t kOffQt = kOffQc + align256(sizeof(double) * 3 * kMaxNP);
constexpr size_t kOffF = kOffQt + align256(sizeof(double) * 3 * kMaxNP);
constexpr size_t kScratchBytes = kOffF + align256(sizeof(double) * 3 * (size_t(kMaxNx) + 1));

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Set-up of one call, written by a kernel from its arguments (no host copy):
// the descriptor with its scratch pointers, the schedule, the order and the
// zeroed work counter.
struct NoAllocSetup {
    Inst1D in;
    int n_print;
    double pt[kMaxPrint];
};

__global__ void k1_noalloc_setup(NoAllocSetup s, unsigned char* base) {
    double* pt = reinterpret_cast<double*>(base + kOffPrint);
    for (int k = threadIdx.x; k < s.n_print; k += blockDim.x) pt[k] = s.pt[k];
    if (threadIdx.x == 0) {
        Inst1D in = s.in;
        in.print_times = pt;
        in.Qc = reinterpret_cast<double*>(base + kOffQc);
        in.Qt = reinterpret_cast<double*>(base + kOffQt);
        in.F = reinterpret_cast<double*>(base + kOffF);
        *reinterpret_cast<Inst1D*>(base + kOffInst) = in;
        int* ord = reinterpret_cast<int*>(base + kOffOrder);
        ord[0] = 0;  // the one instance
        ord[1] = 0;  // work counter
    }
}

constexpr int kMaxDevices = 64;
unsigned char* g_ptr[kMaxDevices];
std::once_flag g_once[kMaxDevices];
hipError_t g_err[kMaxDevices];

// Device address of `scratch` on the current device, taken once per device.
unsigned char* scratch_ptr() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::call_once(g_once[dev], [dev] {
        void* p = nullptr;
        g_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[dev] = g_err[dev] == hipSuccess ? static_cast<unsigned char*>(p) : nullptr;
    });
    return g_ptr[dev];
}

}  // namespace
}  // namespace awcm

extern "C" {

int awcm_run_1d_noalloc(const awcm_instance_1d* p, int flux, const double* print_times, int n_print, double* q,
                        uint8_t* masks, uint8_t* levels, int max_steps, double* t, int* cnt, int* steps_done,
                        long long* updates) {
    using namespace awcm;
    if (!p || !q || !masks || !levels || !t || !cnt || *cnt < 0 || max_steps < 0) return AWCM_EINVAL;
    if (flux != AWCM_FLUX_WENOZ && flux != AWCM_FLUX_WENO5) return AWCM_EINVAL;
    if (n_print < 0 || n_print > kMaxPrint || (n_print > 0 && !print_times)) return AWCM_EINVAL;
    if (p->num_levels < 2 || p->num_levels > 14 || p->nx_coarse < 6) return AWCM_EINVAL;
    if (!(p->dx > 0) || !(p->gamma > 1) || !(p->cfl > 0)) return AWCM_EINVAL;
    const long long nx = (long long)(p->nx_coarse - 1) * (1LL << (p->num_levels - 1)) + 1;
    if (nx > kMaxNx) return AWCM_EINVAL;  // beyond the fixed scratch size
    const int Nx = int(nx), NP = Nx + 2 * G;
    const int nxp = (NP + 15) / 16 * 16;
    const int threads = Nx >= 1024 ? 1024 : 256;
    // the masks and levels stay in LDS: 11 * nxp bytes next to the kernel's own
    int dev = 0, cap = 0;
    hipFuncAttributes fa;
    const void* kfn = flux == AWCM_FLUX_WENO5 ? reinterpret_cast<const void*>(&k1_solve<false, AWCM_FLUX_WENO5, false>)
                                              : reinterpret_cast<const void*>(&k1_solve<false, AWCM_FLUX_WENOZ, false>);
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cap, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        hipFuncGetAttributes(&fa, kfn) != hipSuccess)
        return AWCM_EDEVICE;
    if (size_t(11) * nxp + fa.sharedSizeBytes > size_t(cap)) return AWCM_EINVAL;
    unsigned char* base = scratch_ptr();
    if (!base) return AWCM_EDEVICE;

    NoAllocSetup s;
    std::memset(&s, 0, sizeof(s));
    Inst1D& in = s.in;
    in.L = p->num_levels;
    in.nxc = p->nx_coarse;
    in.Nx = Nx;
    in.n_print = n_print;
    in.gamma = p->gamma;
    in.dx = p->dx;
    in.tol = p->tol;
    in.cfl = p->cfl;
    in.t_final = p->t_final;
    in.Q = q;
    in.masks = masks;
    in.levels = levels;
    in.t = *t;
    in.cnt = *cnt;
    s.n_print = n_print;
    for (int k = 0; k < n_print; k++) s.pt[k] = print_times[k];

    Inst1D* d_inst = reinterpret_cast<Inst1D*>(base + kOffInst);
    int* d_order = reinterpret_cast<int*>(base + kOffOrder);
    hipLaunchKernelGGL(k1_noalloc_setup, dim3(1), dim3(64), 0, 0, s, base);
    if (hipGetLastError() != hipSuccess) return AWCM_EDEVICE;
    const size_t lds = size_t(11) * nxp;
    if (flux == AWCM_FLUX_WENO5)
        hipLaunchKernelGGL((k1_solve<false, AWCM_FLUX_WENO5, false>), dim3(1), dim3(threads), lds, 0,
                           d_inst, 1, d_order, d_order + 1, 1, max_steps, nxp, nullptr);
    else
        hipLaunchKernelGGL((k1_solve<false, AWCM_FLUX_WENOZ, false>), dim3(1), dim3(threads), lds, 0,
                           d_inst, 1, d_order, d_order + 1, 1, max_steps, nxp, nullptr);
    if (hipGetLastError() != hipSuccess) return AWCM_EDEVICE;
    Inst1D out;
    if (hipMemcpy(&out, d_inst, sizeof(out), hipMemcpyDeviceToHost) != hipSuccess) return AWCM_EDEVICE;
    *t = out.t;
    *cnt = out.cnt;
    if (steps_done) *steps_done = out.steps;
    if (updates) *updates = out.sum_active0;
    if (out.status == 0) return AWCM_OK;
    return out.status == -5 ? AWCM_ERANGE : out.status == AWCM_ENONFINITE ? AWCM_ENONFINITE : AWCM_EDEVICE;
}

}  // extern "C"
