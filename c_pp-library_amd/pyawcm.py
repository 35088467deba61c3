"""Python binding of the C ABI in include/awcm_capi.h (libawcm_hip.so).

Loads the in-tree HIP library and raises if it is missing: there is no CPU
fallback behind these classes.  Layouts are the reference's (AoS, ghost 3).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libawcm_hip.so")

AWCM_OK = 0
ERRORS = {-1: "EINVAL", -2: "EDEVICE", -3: "ENOMEM", -4: "ENONFINITE", -5: "ERANGE", -6: "ESTATE", -7: "EDISABLED"}
BC_TRANSMISSIVE = 0
FLUX_WENOZ = 0   # WENOZ_EULER_SOLVER_1D::ComputeFlux pair (the default)
FLUX_WENO5 = 1   # WENO_EULER_SOLVER_1D::ComputedF, the shipped 1-D driver's
BC_DOUBLE_MACH = 1


class AwcmError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} failed: {ERRORS.get(code, code)}")
        self.code = code


class Instance1D(C.Structure):
    _fields_ = [("num_levels", C.c_int), ("nx_coarse", C.c_int), ("gamma", C.c_double), ("dx", C.c_double),
                ("tol", C.c_double), ("cfl", C.c_double), ("t_final", C.c_double)]


_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)
_lib = None


def lib():
    """The loaded HIP library (raises FileNotFoundError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built; run python c_pp-library_amd/build.py")
    L = C.CDLL(LIB_PATH)
    V = C.c_void_p
    I = C.c_int
    D = C.c_double
    sig = {
        "awcm_create_1d": [I, I, D, D, C.POINTER(V)],
        "awcm_create_2d": [I, I, I, D, D, D, I, C.POINTER(V)],
        "awcm_dims": [V, C.POINTER(I), C.POINTER(I)],
        "awcm_upload_q": [V, _dp],
        "awcm_download_q": [V, _dp],
        "awcm_set_mask0": [V, _u8p],
        "awcm_get_masks": [V, _u8p, _i32p],
        "awcm_prepare_euler": [V, D],
        "awcm_interpolate_all_grid": [V, _u8p],
        "awcm_upload_scalar": [V, _dp],
        "awcm_set_flux": [V, I],
        "awcm_batch_set_flux": [V, I],
        "awcm_download_scalar": [V, _dp],
        "awcm_prepare": [V, D],
        "awcm_post_process": [V],
        "awcm_run_scalar": [V, D, D, I],
        "awcm_interpolate": [V, _u8p, _u8p],
        "awcm_interpolate_euler": [V, _u8p, _u8p],
        "awcm_post_process_euler": [V],
        "awcm_set_schedule": [V, _dp, I],
        "awcm_run": [V, D, D, D, I, _dp, C.POINTER(I), C.POINTER(I), C.POINTER(C.c_longlong)],
        "awcm_synchronize": [V],
        "awcm_divide_shared": [_dp, _dp, _dp, C.c_longlong],
        "awcm_batch_create_1d": [I, C.POINTER(Instance1D), _dp, I, C.POINTER(V)],
        "awcm_batch_upload_q": [V, I, _dp],
        "awcm_batch_download_q": [V, I, _dp],
        "awcm_batch_get_masks": [V, I, _u8p, _i32p],
        "awcm_batch_set_launch": [V, I, I],
        "awcm_batch_run": [V, I],
        "awcm_batch_run_sequential": [V, I],
        "awcm_batch_autotune": [V, I, C.POINTER(I), C.POINTER(I), _dp],
        "awcm_batch_state": [V, I, _dp, C.POINTER(I), C.POINTER(I), C.POINTER(C.c_longlong), C.POINTER(I)],
        "awcm_batch_synchronize": [V],
        "awcm_run_timed": [V, D, D, D, I, _dp, C.POINTER(I), C.POINTER(I), C.POINTER(C.c_longlong), _dp],
        "awcm_profile_enable": [V, I],
        "awcm_set_tuning": [V, I, I],
        "awcm_profile_read": [V, I, _dp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)],
        "awcm_add_adjacent_zone": [V, _u8p],
        "awcm_check_reconstruction": [V, _u8p],
        "awcm_add_point_for_differentiation": [V, _u8p, _i32p],
        "awcm_compute_level": [V, _i32p, _u8p],
        "awcm_remove_residue": [V, _dp, I, _u8p],
        "awcm_adjoint_record": [V, D, D, D, I, I, _dp, C.POINTER(I), C.POINTER(I)],
        "awcm_adjoint_replay": [V],
        "awcm_adjoint": [V, _dp, I, _dp],
        "awcm_ic_riemann_1d": [I, D, D, D, _dp, _dp, D, _dp, _dp, _dp],
        "awcm_ic_shu_osher_1d": [I, D, D, D, _dp, D, D, _dp, _dp, _dp],
        "awcm_ic_quadrants_2d": [I, D, _dp, D, _dp, _dp],
        "awcm_ic_double_mach_2d": [I, D, D, _dp, _dp],
        "awcm_ic_box_advection_2d": [I, _dp, _dp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.awcm_destroy.argtypes = [V]
    L.awcm_destroy.restype = None
    L.awcm_batch_destroy.argtypes = [V]
    L.awcm_batch_destroy.restype = None
    _lib = L
    return L


def _check(code, what):
    if code != AWCM_OK:
        raise AwcmError(code, what)


def _d(a):
    return a.ctypes.data_as(_dp)


class Solver:
    """One AWCM solve on the GPU (dim 1 or 2), mirroring AWCM_1D / AWCM_2D."""

    def __init__(self, dim, num_levels, nx_coarse, gamma, dx, ny_coarse=0, x0=0.0, bc_kind=BC_TRANSMISSIVE):
        L = lib()
        self.h = C.c_void_p()
        if dim == 1:
            _check(L.awcm_create_1d(num_levels, nx_coarse, gamma, dx, C.byref(self.h)), "awcm_create_1d")
        else:
            _check(L.awcm_create_2d(num_levels, nx_coarse, ny_coarse, gamma, dx, x0, bc_kind, C.byref(self.h)),
                   "awcm_create_2d")
        self.dim = dim
        nx, ny = C.c_int(), C.c_int()
        _check(L.awcm_dims(self.h, C.byref(nx), C.byref(ny)), "awcm_dims")
        self.nx, self.ny = nx.value, ny.value
        self.t = 0.0
        self.cnt = 1

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().awcm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def npts(self):
        return (self.nx + 6) if self.dim == 1 else (self.nx + 6) * (self.ny + 6)

    @property
    def ncomp(self):
        return 3 if self.dim == 1 else 4

    def upload_q(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.size == self.npts * self.ncomp
        _check(lib().awcm_upload_q(self.h, _d(q)), "awcm_upload_q")

    def download_q(self):
        q = np.empty(self.npts * self.ncomp)
        _check(lib().awcm_download_q(self.h, _d(q)), "awcm_download_q")
        return q

    def set_mask0(self, m0):
        m0 = np.ascontiguousarray(m0, dtype=np.uint8)
        assert m0.size == self.npts
        _check(lib().awcm_set_mask0(self.h, m0.ctypes.data_as(_u8p)), "awcm_set_mask0")

    def masks(self):
        m = np.empty(5 * self.npts, dtype=np.uint8)
        lv = np.empty(4 * self.nx * self.ny, dtype=np.int32)
        _check(lib().awcm_get_masks(self.h, m.ctypes.data_as(_u8p), lv.ctypes.data_as(_i32p)), "awcm_get_masks")
        if self.dim == 1:
            return m.reshape(5, self.nx + 6), lv.reshape(4, self.nx)
        return m.reshape(5, self.ny + 6, self.nx + 6), lv.reshape(4, self.ny, self.nx)

    def prepare(self, tol):
        _check(lib().awcm_prepare_euler(self.h, tol), "awcm_prepare_euler")

    def set_flux(self, kind):
        """1-D stage face fluxes: FLUX_WENOZ or FLUX_WENO5 (awcm_set_flux)."""
        _check(lib().awcm_set_flux(self.h, kind), "awcm_set_flux")

    # ---- the scalar path (2-D): AWCM_2D::Prepare / PostProcess, WAVELET_2D::Interpolate
    def upload_scalar(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.size == self.npts
        _check(lib().awcm_upload_scalar(self.h, _d(q)), "awcm_upload_scalar")

    def download_scalar(self):
        q = np.empty(self.npts)
        _check(lib().awcm_download_scalar(self.h, _d(q)), "awcm_download_scalar")
        return q

    def prepare_scalar(self, tol):
        _check(lib().awcm_prepare(self.h, tol), "awcm_prepare")

    def post_process_scalar(self):
        _check(lib().awcm_post_process(self.h), "awcm_post_process")

    def run_scalar(self, tol, dt, n_steps):
        """n_steps iterations of the box-advection driver (main_2d_advect.cpp) on
        the device (awcm_run_scalar)."""
        _check(lib().awcm_run_scalar(self.h, tol, dt, n_steps), "awcm_run_scalar")

    def interpolate(self, mask_from, mask_to):
        """WAVELET_2D::Interpolate(q, mra, mask_from, mask_to) on the scalar field."""
        f = np.ascontiguousarray(mask_from, dtype=np.uint8)
        t = np.ascontiguousarray(mask_to, dtype=np.uint8)
        assert f.size == self.npts and t.size == self.npts
        _check(lib().awcm_interpolate(self.h, f.ctypes.data_as(_u8p), t.ctypes.data_as(_u8p)), "awcm_interpolate")

    def interpolate_euler(self, mask_from, mask_to):
        """WAVELET_2D::InterpolateEuler(q, mra, mask_from, mask_to)."""
        f = np.ascontiguousarray(mask_from, dtype=np.uint8)
        t = np.ascontiguousarray(mask_to, dtype=np.uint8)
        assert f.size == self.npts and t.size == self.npts
        _check(lib().awcm_interpolate_euler(self.h, f.ctypes.data_as(_u8p), t.ctypes.data_as(_u8p)),
               "awcm_interpolate_euler")

    def interpolate_all_grid(self, mask_from):
        """WAVELET_2D::InterpolateAllGridEuler (Wavelet/WAVELET_2D.h:912-1044) on
        the device field; mask_from is ghosted uint8 in the reference layout."""
        m = np.ascontiguousarray(mask_from, dtype=np.uint8)
        assert m.size == self.npts
        _check(lib().awcm_interpolate_all_grid(self.h, m.ctypes.data_as(_u8p)), "awcm_interpolate_all_grid")

    def post_process(self):
        _check(lib().awcm_post_process_euler(self.h), "awcm_post_process_euler")

    def set_schedule(self, print_times):
        pt = np.ascontiguousarray(print_times, dtype=np.float64)
        _check(lib().awcm_set_schedule(self.h, _d(pt), len(pt)), "awcm_set_schedule")

    def run(self, tol, cfl, t_final, max_steps):
        """Advance up to max_steps steps (stops at t_final); returns (steps, updates)."""
        t = np.array([self.t])
        cnt = C.c_int(self.cnt)
        steps = C.c_int(0)
        upd = C.c_longlong(0)
        _check(lib().awcm_run(self.h, tol, cfl, t_final, max_steps, _d(t), C.byref(cnt), C.byref(steps),
                              C.byref(upd)), "awcm_run")
        self.t = float(t[0])
        self.cnt = cnt.value
        return steps.value, upd.value

    def run_timed(self, tol, cfl, t_final, n_steps):
        """Queue n_steps with no host sync; returns (steps, updates, device ms)."""
        t = np.array([self.t])
        cnt = C.c_int(self.cnt)
        steps = C.c_int(0)
        upd = C.c_longlong(0)
        ms = np.zeros(1)
        _check(lib().awcm_run_timed(self.h, tol, cfl, t_final, n_steps, _d(t), C.byref(cnt), C.byref(steps),
                                    C.byref(upd), _d(ms)), "awcm_run_timed")
        self.t = float(t[0])
        self.cnt = cnt.value
        return steps.value, upd.value, float(ms[0])

    def profile(self, enable=True):
        _check(lib().awcm_profile_enable(self.h, 1 if enable else 0), "awcm_profile_enable")

    def profile_read(self, kernel):
        """(device ms summed over launches, launches, points) of one kernel class."""
        ms = np.zeros(1)
        n, pts = C.c_longlong(), C.c_longlong()
        _check(lib().awcm_profile_read(self.h, kernel, _d(ms), C.byref(n), C.byref(pts)), "awcm_profile_read")
        return float(ms[0]), n.value, pts.value

    def set_tuning(self, param, value):
        _check(lib().awcm_set_tuning(self.h, param, value), "awcm_set_tuning")

    def synchronize(self):
        _check(lib().awcm_synchronize(self.h), "awcm_synchronize")

    # ---- AWCM_MANAGER_1D / AWCM_MANAGER_2D statics on caller arrays ----
    MASK_OPS = ("awcm_add_adjacent_zone", "awcm_check_reconstruction", "awcm_add_point_for_differentiation",
                "awcm_compute_level")

    def mask_op(self, op, mask, level=None):
        """op 0 AddAdjacentZone, 1 CheckReconstruction, 2 AddPointForDifferentiation
        (level read), 3 ComputeLevel (level in/out), on copies; returns (mask, level)."""
        m = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1).copy()
        nl = self.nx * self.ny
        lv = np.ascontiguousarray(level if level is not None else np.zeros(nl), dtype=np.int32).reshape(-1).copy()
        L = lib()
        mp, lp = m.ctypes.data_as(_u8p), lv.ctypes.data_as(_i32p)
        if op == 0:
            st = L.awcm_add_adjacent_zone(self.h, mp)
        elif op == 1:
            st = L.awcm_check_reconstruction(self.h, mp)
        elif op == 2:
            st = L.awcm_add_point_for_differentiation(self.h, mp, lp)
        else:
            st = L.awcm_compute_level(self.h, lp, mp)
        _check(st, self.MASK_OPS[op])
        return m, lv

    def remove_residue(self, q, ncomp, mask):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1).copy()
        m = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
        _check(lib().awcm_remove_residue(self.h, _d(q), ncomp, m.ctypes.data_as(_u8p)), "awcm_remove_residue")
        return q

    # ---- discrete adjoint (1-D) ----
    def adjoint_record(self, tol, cfl, t_final, max_steps, checkpoint_every=32):
        """awcm_run that also records the frozen trajectory; returns steps."""
        t = np.array([self.t])
        cnt = C.c_int(self.cnt)
        steps = C.c_int(0)
        _check(lib().awcm_adjoint_record(self.h, tol, cfl, t_final, max_steps, checkpoint_every, _d(t),
                                         C.byref(cnt), C.byref(steps)), "awcm_adjoint_record")
        self.t = float(t[0])
        self.cnt = cnt.value
        return steps.value

    def adjoint_replay(self):
        """Frozen forward over the recorded steps, from the current state."""
        _check(lib().awcm_adjoint_replay(self.h), "awcm_adjoint_replay")

    def adjoint(self, dJ_dQfinal, checkpoint_every=0):
        """dJ/dQ0 through the recorded steps (layout of upload_q)."""
        g = np.ascontiguousarray(dJ_dQfinal, dtype=np.float64)
        assert g.size == self.npts * self.ncomp
        out = np.empty_like(g)
        _check(lib().awcm_adjoint(self.h, _d(g), checkpoint_every, _d(out)), "awcm_adjoint")
        return out


KERNEL_STAGE_WENO, KERNEL_STAGE_COARSE, KERNEL_PREPARE = 0, 1, 2
TUNE_TAIL_CELLS, TUNE_GRAPH = 0, 1

# 15-digit four-quadrant states (rho, u, v, p), quadrants NE, NW, SW, SE
# (PROBLEM_QUADRANTS.h:43-73)
QUAD_STATES = np.array([
    [1.5, 0.0, 0.0, 1.5],
    [0.532258064516129, 1.206045378311055, 0.0, 0.3],
    [0.137992831541219, 1.206045378311055, 1.206045378311055, 0.029032258064516],
    [0.532258064516129, 0.0, 1.206045378311055, 0.3],
], dtype=np.float64)


def ic_riemann_1d(res, xl, xu, split, left, right, gamma=1.4):
    """SHOCKTUBE::Initialize on [xl, xu]; returns (q, dx, x0)."""
    q = np.empty((res + 6) * 3)
    a = np.ascontiguousarray(left, dtype=np.float64)
    b = np.ascontiguousarray(right, dtype=np.float64)
    dx, x0 = np.zeros(1), np.zeros(1)
    _check(lib().awcm_ic_riemann_1d(res, xl, xu, split, _d(a), _d(b), gamma, _d(q), _d(dx), _d(x0)), "awcm_ic_riemann_1d")
    return q, float(dx[0]), float(x0[0])


def ic_shu_osher_1d(res, xl=-5.0, xu=5.0, split=-4.0, left=(3.857143, 2.629369, 10.33333), amp=0.2, gamma=1.4):
    """ENTROPY_INTERACTING::Initialize; returns (q, dx, x0)."""
    q = np.empty((res + 6) * 3)
    a = np.ascontiguousarray(left, dtype=np.float64)
    dx, x0 = np.zeros(1), np.zeros(1)
    _check(lib().awcm_ic_shu_osher_1d(res, xl, xu, split, _d(a), amp, gamma, _d(q), _d(dx), _d(x0)),
           "awcm_ic_shu_osher_1d")
    return q, float(dx[0]), float(x0[0])


def ic_quadrants_2d(res, split=0.5, states=QUAD_STATES, gamma=1.4):
    """QUADRANTS::Initialize on [0,1]^2; returns (q, dx)."""
    q = np.empty((res + 6) * (res + 6) * 4)
    st = np.ascontiguousarray(states, dtype=np.float64)
    dx = np.zeros(1)
    _check(lib().awcm_ic_quadrants_2d(res, split, _d(st), gamma, _d(q), _d(dx)), "awcm_ic_quadrants_2d")
    return q, float(dx[0])


def ic_double_mach_2d(res, rho_pre=1.4, gamma=1.4):
    """DOUBLEMACH::Initialize, ((res-1)*4+1) x res points; returns (q, dx)."""
    nx = (res - 1) * 4 + 1
    q = np.empty((nx + 6) * (res + 6) * 4)
    dx = np.zeros(1)
    _check(lib().awcm_ic_double_mach_2d(res, rho_pre, gamma, _d(q), _d(dx)), "awcm_ic_double_mach_2d")
    return q, float(dx[0])


def ic_box_advection_2d(res):
    """BOX_ADVECTION::Initialize: the scalar field ((res+6)^2 doubles) and dx."""
    q = np.empty((res + 6) * (res + 6))
    dx = np.zeros(1)
    _check(lib().awcm_ic_box_advection_2d(res, _d(q), _d(dx)), "awcm_ic_box_advection_2d")
    return q, float(dx[0])


def divide_shared(a, b):
    """a / b elementwise on the device, as the kernels divide by a shared divisor."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    if a.shape != b.shape:
        raise ValueError("a and b differ in shape")
    q = np.empty_like(a)
    _check(lib().awcm_divide_shared(_d(a), _d(b), _d(q), a.size), "awcm_divide_shared")
    return q


def print_times(t_final, num_print):
    """arr_printing_times of the drivers (main_1d_euler.cpp:43-48)."""
    inc = t_final / (num_print - 1)
    return np.array([inc * p for p in range(num_print)], dtype=np.float64)


class Batch1D:
    """Many independent 1-D solves in one persistent launch."""

    def __init__(self, params, print_times):
        arr = (Instance1D * len(params))()
        for k, p in enumerate(params):
            arr[k] = Instance1D(*p)
        pt = np.ascontiguousarray(print_times, dtype=np.float64)
        self.h = C.c_void_p()
        _check(lib().awcm_batch_create_1d(len(params), arr, _d(pt), len(pt), C.byref(self.h)), "awcm_batch_create_1d")
        self.n = len(params)
        self.nx = [(p[1] - 1) * 2 ** (p[0] - 1) + 1 for p in params]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().awcm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_flux(self, kind):
        """Stage face fluxes of every instance: FLUX_WENOZ or FLUX_WENO5."""
        _check(lib().awcm_batch_set_flux(self.h, kind), "awcm_batch_set_flux")

    def upload_q(self, idx, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.size == (self.nx[idx] + 6) * 3
        _check(lib().awcm_batch_upload_q(self.h, idx, _d(q)), "awcm_batch_upload_q")

    def download_q(self, idx):
        q = np.empty((self.nx[idx] + 6) * 3)
        _check(lib().awcm_batch_download_q(self.h, idx, _d(q)), "awcm_batch_download_q")
        return q

    def masks(self, idx):
        n = self.nx[idx]
        m = np.empty(5 * (n + 6), dtype=np.uint8)
        lv = np.empty(4 * n, dtype=np.int32)
        _check(lib().awcm_batch_get_masks(self.h, idx, m.ctypes.data_as(_u8p), lv.ctypes.data_as(_i32p)),
               "awcm_batch_get_masks")
        return m.reshape(5, n + 6), lv.reshape(4, n)

    def set_launch(self, threads, workgroups=0):
        _check(lib().awcm_batch_set_launch(self.h, threads, workgroups), "awcm_batch_set_launch")

    def run(self, max_steps=1 << 30):
        _check(lib().awcm_batch_run(self.h, max_steps), "awcm_batch_run")

    def run_sequential(self, max_steps=1 << 30):
        _check(lib().awcm_batch_run_sequential(self.h, max_steps), "awcm_batch_run_sequential")

    def autotune(self, probe_steps=20):
        """Pick the batched launch shape by timing probes; returns (threads, workgroups, probe_ms[9])."""
        t, w = C.c_int(), C.c_int()
        ms = np.zeros(9)
        _check(lib().awcm_batch_autotune(self.h, probe_steps, C.byref(t), C.byref(w), _d(ms)), "awcm_batch_autotune")
        return t.value, w.value, ms

    def state(self, idx):
        t = np.zeros(1)
        cnt, steps, status = C.c_int(), C.c_int(), C.c_int()
        upd = C.c_longlong()
        _check(lib().awcm_batch_state(self.h, idx, _d(t), C.byref(cnt), C.byref(steps), C.byref(upd),
                                      C.byref(status)), "awcm_batch_state")
        return dict(t=float(t[0]), cnt=cnt.value, steps=steps.value, updates=upd.value, status=status.value)
