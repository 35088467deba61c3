"""The 1-D solve without allocation (awcm_run_1d_noalloc) against awcm_run.

awcm_run_1d_noalloc runs awcm_run's 1-D time loop on caller-owned device
buffers, with the library's working memory in one fixed device array per GPU.
Two calls in a row -- a 1025-point Sod run, then a smaller 513-point
Shu-Osher run that the first call's leftovers in that array must not touch --
each give the state, masks, levels, t, print counter and step count of the
same run through a solver handle (awcm_run), bit for bit.  The test's device
buffers come from hipMalloc of the HIP runtime the library itself uses, so the
process holds one HIP runtime.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the library's HIP runtime (test buffers)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _setup():
    import pyawcm as P
    L = P.lib()
    f = L.awcm_run_1d_noalloc
    f.restype = C.c_int
    f.argtypes = [C.POINTER(P.Instance1D), C.c_int, C.POINTER(C.c_double), C.c_int, C.c_void_p, C.c_void_p,
                  C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int),
                  C.POINTER(C.c_longlong)]
    return Hip(), P, f


def _handle_run(P, L, q, dx, tol, cfl, tf, pts, flux, steps):
    s = P.Solver(1, L, 17, 1.4, dx)
    if flux:
        s.set_flux(flux)
    s.set_schedule(pts)
    s.upload_q(q)
    n, _ = s.run(tol, cfl, tf, steps)
    m, lv = s.masks()
    return s.download_q(), m, lv, s.t, s.cnt, n


def _noalloc_run(hip, P, f, L, q, dx, tol, cfl, tf, pts, flux, steps):
    nx = 16 * 2 ** (L - 1) + 1
    NP = nx + 6
    soa = np.ascontiguousarray(q.reshape(NP, 3).T)  # 3 planes of Nx+6 doubles
    masks = np.zeros((5, NP), np.uint8)
    masks[1] = 1  # AWCM_1D(mra): mask0 all true, the others false, levels 0
    levels = np.zeros(4 * nx, np.uint8)
    dq, dm, dl = hip.upload(soa), hip.upload(masks), hip.upload(levels)
    p = P.Instance1D(L, 17, 1.4, dx, tol, cfl, tf)
    t, cnt, n, upd = C.c_double(0.0), C.c_int(1), C.c_int(0), C.c_longlong(0)
    pt = np.ascontiguousarray(pts, dtype=np.float64)
    st = f(C.byref(p), flux, pt.ctypes.data_as(C.POINTER(C.c_double)), len(pt), dq, dm, dl, steps, C.byref(t),
           C.byref(cnt), C.byref(n), C.byref(upd))
    assert st == 0, st
    qo = hip.download(dq, soa).T.reshape(-1).copy()
    mo = hip.download(dm, masks).reshape(-1)
    lo = hip.download(dl, levels).astype(np.int32)
    hip.free()
    return qo, mo, lo, t.value, cnt.value, n.value


CASES = [
    # (levels, problem, tol, cfl, t_final, n_print, flux, max steps)
    (7, "sod", 1e-4, 0.5, 0.2, 41, 0, 120),
    (6, "shu", 1e-4, 0.5, 0.1, 2, 0, 60),
    (5, "sod", 1e-3, 0.5, 0.2, 41, 1, 40),  # WENO5-JS stage fluxes
]


def _ic(L, kind):
    nx = 16 * 2 ** (L - 1) + 1
    if kind == "sod":
        q, dx, _ = O.ic1(0, nx, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    else:
        q, dx, _ = O.ic1(1, nx, -5.0, 5.0, -4.0, O.SHU_L, O.SHU_L)
    return q, dx


def test_noalloc_matches_awcm_run_twice():
    hip, P, f = _setup()
    for L, kind, tol, cfl, tf, npr, flux, steps in CASES:
        print("no-alloc case", L, kind, flush=True)
        q, dx = _ic(L, kind)
        pts = O.print_times(tf, npr)
        qa, ma, la, ta, ca, na = _handle_run(P, L, q, dx, tol, cfl, tf, pts, flux, steps)
        qb, mb, lb, tb, cb, nb = _noalloc_run(hip, P, f, L, q, dx, tol, cfl, tf, pts, flux, steps)
        assert na == nb and na > 30
        assert ta == tb and ca == cb
        assert np.array_equal(ma.reshape(-1), mb) and np.array_equal(la.reshape(-1), lb)
        assert np.array_equal(qa.view(np.uint64), qb.view(np.uint64))


def test_noalloc_rejects_beyond_its_limits():
    hip, P, f = _setup()
    q = hip.upload(np.zeros(8))
    t, cnt = C.c_double(0.0), C.c_int(1)
    big = P.Instance1D(11, 17, 1.4, 1e-4, 1e-4, 0.5, 0.1)  # 16385 points > 8193
    pt = np.zeros(65)
    ptr = pt.ctypes.data_as(C.POINTER(C.c_double))
    assert f(C.byref(big), 0, ptr, 1, q, q, q, 1, C.byref(t), C.byref(cnt), None, None) == -1
    ok = P.Instance1D(6, 17, 1.4, 1e-3, 1e-4, 0.5, 0.1)
    assert f(C.byref(ok), 0, ptr, 65, q, q, q, 1, C.byref(t), C.byref(cnt), None, None) == -1  # > 64 output times
    hip.free()
