"""ctypes bindings for the CPU oracle (oracle/liboracle_awcm.so) and for the
reference build (oracle/_ref/libawcm_ref.so), which export the same entry
points under the prefixes ``orc`` and ``ref``.

Test infrastructure only: the product never imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle_awcm.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libawcm_ref.so")

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)
_ip = C.POINTER(C.c_int)


def dptr(a):
    return a.ctypes.data_as(_dp)


def u8ptr(a):
    return a.ctypes.data_as(_u8p)


def i32ptr(a):
    return a.ctypes.data_as(_i32p)


_libs = {}


def load(prefix="orc"):
    """Load the oracle (prefix 'orc') or the reference build (prefix 'ref')."""
    if prefix in _libs:
        return _libs[prefix]
    path = ORACLE_SO if prefix == "orc" else REF_SO
    if not os.path.exists(path):
        raise FileNotFoundError(path + " is not built (make -C oracle)")
    lib = C.CDLL(path)
    p = prefix
    V = C.c_void_p
    sig = {
        f"{p}1_create": (V, [C.c_int, C.c_int, C.c_double, C.c_double]),
        f"{p}1_destroy": (None, [V]),
        f"{p}1_nx": (C.c_int, [V]),
        f"{p}1_set_q": (None, [V, _dp]),
        f"{p}1_get_q": (None, [V, _dp]),
        f"{p}1_get_masks": (None, [V, _u8p, _i32p]),
        f"{p}1_set_mask0": (None, [V, _u8p]),
        f"{p}1_prepare": (None, [V, C.c_double]),
        f"{p}1_set_flux": (None, [V, C.c_int]),
        f"{p}1_set_scalar": (None, [V, _dp]),
        f"{p}1_get_scalar": (None, [V, _dp]),
        f"{p}1_prepare_scalar": (None, [V, C.c_double]),
        f"{p}1_interpolate_scalar": (None, [V, _u8p, _u8p]),
        f"{p}1_interpolate_euler": (None, [V, _u8p, _u8p]),
        f"{p}_flux1_js": (None, [C.c_double, _dp, C.c_double, _dp]),
        f"{p}1_step": (C.c_double, [V, C.c_double, C.c_double, _dp, C.c_int, _ip, _dp]),
        f"{p}2_nx": (C.c_int, [V]),
        f"{p}2_ny": (C.c_int, [V]),
        f"{p}2_destroy": (None, [V]),
        f"{p}2_set_q": (None, [V, _dp]),
        f"{p}2_get_q": (None, [V, _dp]),
        f"{p}2_get_masks": (None, [V, _u8p, _i32p]),
        f"{p}2_set_mask0": (None, [V, _u8p]),
        f"{p}2_prepare": (None, [V, C.c_double]),
        f"{p}2_interpolate_all": (None, [V, _u8p]),
        f"{p}2_interpolate_euler": (None, [V, _u8p, _u8p]),
        f"{p}2_set_scalar": (None, [V, _dp]),
        f"{p}2_get_scalar": (None, [V, _dp]),
        f"{p}2_prepare_scalar": (None, [V, C.c_double]),
        f"{p}2_interpolate_scalar": (None, [V, _u8p, _u8p]),
        f"{p}2_step": (C.c_double, [V, C.c_double, C.c_double, _dp, C.c_int, _ip, _dp]),
        f"{p}1_mask_op": (None, [V, C.c_int, _u8p, _i32p]),
        f"{p}2_mask_op": (None, [V, C.c_int, _u8p, _i32p]),
        f"{p}1_remove_residue": (None, [V, _dp, C.c_int, _u8p]),
        f"{p}2_remove_residue": (None, [V, _dp, C.c_int, _u8p]),
        f"{p}_flux1": (None, [C.c_double, _dp, C.c_double, _dp]),
        f"{p}_flux2": (None, [C.c_double, C.c_int, _dp, C.c_double, _dp]),
        f"{p}_wenoz_left": (C.c_double, [C.c_double] * 5),
        f"{p}_wenoz_right": (C.c_double, [C.c_double] * 5),
        f"{p}_weights": (None, [_dp]),
    }
    if prefix == "orc":
        sig["orc2_create"] = (V, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int])
        sig["orc_ic1"] = (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                    C.c_double, C.c_double, _dp, _dp, _dp])
        sig["orc_ic2"] = (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_double, _dp, C.c_double, _dp, _dp, _dp])
        sig["orc1_oob"] = (C.c_long, [V])
        sig["orc1_record"] = (None, [V, C.c_int])
        sig["orc1_tape_steps"] = (C.c_int, [V])
        sig["orc1_tape_scalars"] = (None, [V, _dp])
        sig["orc1_replay"] = (None, [V, C.c_int, C.c_int])
        sig["orc2_oob"] = (C.c_long, [V])
        sig["orc2_advect"] = (None, [V, C.c_double, C.c_double, C.c_int, C.c_int])
        sig["orc_fnv1a"] = (C.c_uint64, [_u8p, C.c_longlong, C.c_uint64])
    else:
        sig["ref2_create"] = (V, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_int])
        sig["ref_ic1"] = (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp])
        sig["ref_ic2"] = (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp])
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _libs[prefix] = lib
    return lib


def have_ref():
    return os.path.exists(REF_SO)


# ---------------------------------------------------------------------------
# Problem set-ups (own restatement; see oracle/awcm_oracle.cpp orc_ic1/orc_ic2)
SOD_L = (1.0, 0.0, 1.0)
SOD_R = (0.125, 0.0, 0.1)
SHU_L = (3.857143, 2.629369, 10.33333)
# 15-digit four-quadrant states (rho, u, v, p), quadrants 1..4 (PROBLEM_QUADRANTS.h:43-73)
QUAD_STATES = np.array([
    [1.5, 0.0, 0.0, 1.5],
    [0.532258064516129, 1.206045378311055, 0.0, 0.3],
    [0.137992831541219, 1.206045378311055, 1.206045378311055, 0.029032258064516],
    [0.532258064516129, 0.0, 1.206045378311055, 0.3],
], dtype=np.float64)


def ic1(kind, res, xl, xu, split, sl, sr, amp=0.2, gamma=1.4):
    lib = load("orc")
    q = np.zeros((res + 6) * 3)
    dx = np.zeros(1)
    x0 = np.zeros(1)
    a = np.array(sl, dtype=np.float64)
    b = np.array(sr, dtype=np.float64)
    lib.orc_ic1(kind, res, xl, xu, split, dptr(a), dptr(b), amp, gamma, dptr(q), dptr(dx), dptr(x0))
    return q, float(dx[0]), float(x0[0])


def ic2(kind, nx, ny, xl, xu, yl, yu, s, states=None, gamma=1.4):
    lib = load("orc")
    q = np.zeros((nx + 6) * (ny + 6) * 4)
    dx = np.zeros(1)
    dy = np.zeros(1)
    st = np.ascontiguousarray(states if states is not None else QUAD_STATES, dtype=np.float64)
    lib.orc_ic2(kind, nx, ny, xl, xu, yl, yu, s, dptr(st), gamma, dptr(q), dptr(dx), dptr(dy))
    return q, float(dx[0]), float(dy[0])


def box_grid(res):
    """BOX_ADVECTION's cell-centre grid (PROBLEM_BOX_ADVECTION.h:26-33): the
    res+1 face grid on [0, 1] gives dx' = 1/res; the cell grid runs from dx'/2
    to 1 - dx'/2 in res points.  Returns (x_min, x_max, dx)."""
    dxx = (1.0 - 0.0) / ((res + 1) - 1)
    xl, xu = 0.0 + 0.5 * dxx, 1.0 - 0.5 * dxx
    return xl, xu, (xu - xl) / (res - 1)


def ic_box(res):
    """The scalar box-advection field, (res+6)^2 doubles, and its dx."""
    xl, xu, _ = box_grid(res)
    q, dx, _ = ic2(2, res, res, xl, xu, xl, xu, 0.0)
    return q.reshape(-1, 4)[:, 0].copy(), dx


FNV_OFFSET = 1469598103934665603


def fnv1a(data, h=FNV_OFFSET):
    """FNV-1a (64-bit) of the bytes of `data`, continuing from h."""
    b = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    return int(load("orc").orc_fnv1a(u8ptr(b), b.size, h))


def print_times(t_final, num_print):
    """arr_printing_times of the drivers (main_1d_euler.cpp:43-48)."""
    inc = t_final / (num_print - 1)
    return np.array([inc * p for p in range(num_print)], dtype=np.float64)


class Solver1D:
    """One 1-D AWCM solve on the oracle ('orc') or the reference build ('ref')."""

    def __init__(self, prefix, num_levels, nx_coarse, gamma, dx):
        self.lib = load(prefix)
        self.p = prefix
        self.h = getattr(self.lib, f"{prefix}1_create")(num_levels, nx_coarse, gamma, dx)
        self.nx = getattr(self.lib, f"{prefix}1_nx")(self.h)
        self.t = np.zeros(1)
        self.cnt = C.c_int(1)

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.lib, f"{self.p}1_destroy")(self.h)
            self.h = None

    def set_q(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.size == (self.nx + 6) * 3
        getattr(self.lib, f"{self.p}1_set_q")(self.h, dptr(q))

    def get_q(self):
        q = np.zeros((self.nx + 6) * 3)
        getattr(self.lib, f"{self.p}1_get_q")(self.h, dptr(q))
        return q

    def set_mask0(self, m0):
        m0 = np.ascontiguousarray(m0, dtype=np.uint8)
        getattr(self.lib, f"{self.p}1_set_mask0")(self.h, u8ptr(m0))

    def masks(self):
        m = np.zeros(5 * (self.nx + 6), dtype=np.uint8)
        lv = np.zeros(4 * self.nx, dtype=np.int32)
        getattr(self.lib, f"{self.p}1_get_masks")(self.h, u8ptr(m), i32ptr(lv))
        return m.reshape(5, self.nx + 6), lv.reshape(4, self.nx)

    def prepare(self, tol):
        getattr(self.lib, f"{self.p}1_prepare")(self.h, tol)

    def mask_op(self, op, mask, level=None):
        """AWCM_MANAGER_1D statics (op 0 AddAdjacentZone, 1 CheckReconstruction,
        2 AddPointForDifferentiation, 3 ComputeLevel) on copies; returns (mask, level)."""
        m = np.ascontiguousarray(mask, dtype=np.uint8).copy()
        lv = np.ascontiguousarray(level if level is not None else np.zeros(self.nx), dtype=np.int32).copy()
        getattr(self.lib, f"{self.p}1_mask_op")(self.h, op, u8ptr(m), i32ptr(lv))
        return m, lv

    def remove_residue(self, q, ncomp, mask):
        q = np.ascontiguousarray(q, dtype=np.float64).copy()
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        getattr(self.lib, f"{self.p}1_remove_residue")(self.h, dptr(q), ncomp, u8ptr(m))
        return q

    def set_scalar(self, q):
        """The scalar field of AWCM_1D::Prepare ((nx+6) doubles)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        getattr(self.lib, f"{self.p}1_set_scalar")(self.h, dptr(q))

    def get_scalar(self):
        q = np.zeros(self.nx + 6)
        getattr(self.lib, f"{self.p}1_get_scalar")(self.h, dptr(q))
        return q

    def prepare_scalar(self, tol):
        getattr(self.lib, f"{self.p}1_prepare_scalar")(self.h, tol)

    def interpolate_scalar(self, mask_from, mask_to=None):
        """WAVELET_1D::Interpolate(q, mra, from[, to]); no `to`: every point."""
        f = np.ascontiguousarray(mask_from, dtype=np.uint8)
        t = None if mask_to is None else np.ascontiguousarray(mask_to, dtype=np.uint8)
        getattr(self.lib, f"{self.p}1_interpolate_scalar")(self.h, u8ptr(f), None if t is None else u8ptr(t))

    def interpolate_euler(self, mask_from, mask_to=None):
        """WAVELET_1D::InterpolateEuler(q, mra, from[, to]); no `to`: every
        point, all three components."""
        f = np.ascontiguousarray(mask_from, dtype=np.uint8)
        t = None if mask_to is None else np.ascontiguousarray(mask_to, dtype=np.uint8)
        getattr(self.lib, f"{self.p}1_interpolate_euler")(self.h, u8ptr(f), None if t is None else u8ptr(t))

    def set_flux(self, kind):
        """Face fluxes of the stages: 0 the WENO-Z pair, 1 the shipped driver's
        WENO5-JS ComputedF (main_1d_euler.cpp:109)."""
        getattr(self.lib, f"{self.p}1_set_flux")(self.h, kind)

    def step(self, tol, cfl, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        return getattr(self.lib, f"{self.p}1_step")(self.h, tol, cfl, dptr(pts), len(pts),
                                                       C.byref(self.cnt), dptr(self.t))

    # frozen-trajectory tape (oracle only)
    def record(self, on=True):
        self.lib.orc1_record(self.h, 1 if on else 0)

    def tape_steps(self):
        return self.lib.orc1_tape_steps(self.h)

    def tape_scalars(self):
        """(steps, 4) array: dt and the three stage alphas of each recorded step."""
        n = self.tape_steps()
        out = np.zeros(4 * max(n, 1))
        self.lib.orc1_tape_scalars(self.h, dptr(out))
        return out[:4 * n].reshape(n, 4)

    def replay(self, first=0, count=None):
        """Replay recorded steps from the current state with everything frozen."""
        count = self.tape_steps() - first if count is None else count
        self.lib.orc1_replay(self.h, first, count)


class Solver2D:
    """One 2-D AWCM solve on the oracle ('orc') or the reference build ('ref').

    bc_kind 0: zero-order extrapolation (quadrants); 1: double Mach reflection.
    """

    def __init__(self, prefix, num_levels, nxc, nyc, gamma, bc_kind, xl, xu, yl, yu):
        self.lib = load(prefix)
        self.p = prefix
        if prefix == "orc":
            nx = (nxc - 1) * 2 ** (num_levels - 1) + 1
            dx = (xu - xl) / (nx - 1)
            self.h = self.lib.orc2_create(num_levels, nxc, nyc, gamma, dx, xl, bc_kind)
        else:
            self.h = self.lib.ref2_create(num_levels, nxc, nyc, gamma, xl, xu, yl, yu, bc_kind)
        self.nx = getattr(self.lib, f"{prefix}2_nx")(self.h)
        self.ny = getattr(self.lib, f"{prefix}2_ny")(self.h)
        self.t = np.zeros(1)
        self.cnt = C.c_int(1)

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.lib, f"{self.p}2_destroy")(self.h)
            self.h = None

    @property
    def npts(self):
        return (self.nx + 6) * (self.ny + 6)

    def set_q(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.size == self.npts * 4
        getattr(self.lib, f"{self.p}2_set_q")(self.h, dptr(q))

    def get_q(self):
        q = np.zeros(self.npts * 4)
        getattr(self.lib, f"{self.p}2_get_q")(self.h, dptr(q))
        return q

    def set_mask0(self, m0):
        m0 = np.ascontiguousarray(m0, dtype=np.uint8)
        getattr(self.lib, f"{self.p}2_set_mask0")(self.h, u8ptr(m0))

    def masks(self):
        m = np.zeros(5 * self.npts, dtype=np.uint8)
        lv = np.zeros(4 * self.nx * self.ny, dtype=np.int32)
        getattr(self.lib, f"{self.p}2_get_masks")(self.h, u8ptr(m), i32ptr(lv))
        return m.reshape(5, self.ny + 6, self.nx + 6), lv.reshape(4, self.ny, self.nx)

    def prepare(self, tol):
        getattr(self.lib, f"{self.p}2_prepare")(self.h, tol)

    def mask_op(self, op, mask, level=None):
        """AWCM_MANAGER_2D statics (op 0 AddAdjacentZone, 1 CheckReconstruction,
        2 AddPointForDifferentiation, 3 ComputeLevel) on copies; returns (mask, level)."""
        m = np.ascontiguousarray(mask, dtype=np.uint8).copy()
        lv = np.ascontiguousarray(level if level is not None else np.zeros(self.nx * self.ny), dtype=np.int32).copy()
        getattr(self.lib, f"{self.p}2_mask_op")(self.h, op, u8ptr(m), i32ptr(lv))
        return m, lv

    def remove_residue(self, q, ncomp, mask):
        q = np.ascontiguousarray(q, dtype=np.float64).copy()
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        getattr(self.lib, f"{self.p}2_remove_residue")(self.h, dptr(q), ncomp, u8ptr(m))
        return q

    def set_scalar(self, q):
        """The scalar field of AWCM_2D::Prepare ((ny+6)*(nx+6) doubles)."""
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        assert q.size == self.npts
        getattr(self.lib, f"{self.p}2_set_scalar")(self.h, dptr(q))

    def get_scalar(self):
        q = np.zeros(self.npts)
        getattr(self.lib, f"{self.p}2_get_scalar")(self.h, dptr(q))
        return q

    def prepare_scalar(self, tol):
        """AWCM_2D::Prepare on the scalar field."""
        getattr(self.lib, f"{self.p}2_prepare_scalar")(self.h, tol)

    def advect(self, tol, dt, n=1, with_prepare=True):
        """n iterations of the box-advection driver (Wavelet/main_2d_advect.cpp)
        on the scalar field; with_prepare=False runs one iteration's stages,
        combine and PostProcess without its Prepare."""
        self.lib.orc2_advect(self.h, tol, dt, n, 1 if with_prepare else 0)

    def interpolate_scalar(self, mask_from, mask_to):
        """WAVELET_2D::Interpolate(q, mra, mask_from, mask_to) on the scalar field."""
        f = np.ascontiguousarray(mask_from, dtype=np.uint8).reshape(-1)
        t = np.ascontiguousarray(mask_to, dtype=np.uint8).reshape(-1)
        assert f.size == self.npts and t.size == self.npts
        getattr(self.lib, f"{self.p}2_interpolate_scalar")(self.h, u8ptr(f), u8ptr(t))

    def interpolate_euler(self, mask_from, mask_to):
        """WAVELET_2D::InterpolateEuler(q, mra, mask_from, mask_to)."""
        f = np.ascontiguousarray(mask_from, dtype=np.uint8).reshape(-1)
        t = np.ascontiguousarray(mask_to, dtype=np.uint8).reshape(-1)
        assert f.size == self.npts and t.size == self.npts
        getattr(self.lib, f"{self.p}2_interpolate_euler")(self.h, u8ptr(f), u8ptr(t))

    def interpolate_all(self, mask_from):
        """WAVELET_2D::InterpolateAllGridEuler with mask_from ((ny+6, nx+6) bytes)."""
        m = np.ascontiguousarray(mask_from, dtype=np.uint8).reshape(-1)
        assert m.size == self.npts
        getattr(self.lib, f"{self.p}2_interpolate_all")(self.h, u8ptr(m))

    def step(self, tol, cfl, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        return getattr(self.lib, f"{self.p}2_step")(self.h, tol, cfl, dptr(pts), len(pts),
                                                       C.byref(self.cnt), dptr(self.t))
