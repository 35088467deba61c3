"""The headline workload at its full size, through properties that hold at any
size: the double Mach reflection on 4097 x 1025 finest points (7 levels,
65 x 17 coarse, eps 1e-3, CFL 0.4, pre-shock density 1.4) run from t = 0 to
t = 0.2 twice, once as graph replays of the whole step and once with every
kernel launched on its own.

* Both runs end at t = 0.2 exactly, after the same number of steps, with the
  same field, mask and level bits: the step has no arithmetic whose result
  depends on the order work is done in (maxima are taken over bit patterns,
  work-list order does not enter any value).
* Every field value is finite; density and pressure are positive at every
  point of mask0, and every interior point outside mask0 holds zero (the
  residue removal that ends each step).
* The masks are nested (mask0 within mask1 within mask2 within mask3) and
  every level lies in 0..6.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L, NXC, NYC, TOL, CFL, T_FINAL = 7, 65, 17, 1e-3, 0.4, 0.2


def _solver(P, eager):
    ny = (NYC - 1) * 2 ** (L - 1) + 1
    q, dx = P.ic_double_mach_2d(ny, 1.4)
    old = os.environ.get("AWCM_NO_GRAPH")
    if eager:
        os.environ["AWCM_NO_GRAPH"] = "1"  # read when the solver is created
    try:
        S = P.Solver(2, L, NXC, 1.4, dx, ny_coarse=NYC, x0=0.0, bc_kind=P.BC_DOUBLE_MACH)
    finally:
        if old is None:
            os.environ.pop("AWCM_NO_GRAPH", None)
        else:
            os.environ["AWCM_NO_GRAPH"] = old
    S.set_schedule(P.print_times(T_FINAL, 21))
    S.upload_q(q)
    return S


def _run(S, what):
    steps = 0
    while S.t < T_FINAL:
        n, _ = S.run(TOL, CFL, T_FINAL, 2048)
        assert n > 0
        steps += n
        print(f"{what}: t={S.t:.5f} steps={steps}", flush=True)
    return steps


def test_dmr_full_run_graph_and_eager_agree():
    import pyawcm as P
    P.lib()
    runs = []
    for eager in (False, True):
        S = _solver(P, eager)
        steps = _run(S, "eager" if eager else "graph")
        m, lv = S.masks()
        runs.append((steps, S.t, S.cnt, S.download_q(), m, lv))
        del S
    (s0, t0, c0, q0, m0, l0), (s1, t1, c1, q1, m1, l1) = runs
    assert t0 == t1 == T_FINAL and s0 == s1 and c0 == c1
    assert s0 > 10000
    assert np.array_equal(q0.view(np.uint64), q1.view(np.uint64)), "fields differ between graph and eager runs"
    assert np.array_equal(m0, m1) and np.array_equal(l0, l1), "masks or levels differ between graph and eager runs"

    nx = (NXC - 1) * 2 ** (L - 1) + 1
    ny = (NYC - 1) * 2 ** (L - 1) + 1
    q = q0.reshape(ny + 6, nx + 6, 4)
    assert np.isfinite(q).all()
    mask0 = m0[1] != 0
    inner = np.zeros_like(mask0)
    inner[3:-3, 3:-3] = True
    rho, mx, my, E = (q[..., k] for k in range(4))
    act = mask0 & inner
    p = 0.4 * (E[act] - 0.5 * (mx[act] ** 2 + my[act] ** 2) / rho[act])
    assert (rho[act] > 0).all() and (p > 0).all()
    assert not q[inner & ~mask0].any(), "non-zero state at interior points outside mask0"
    for k in range(1, 4):
        assert not (m0[k] & ~m0[k + 1]).any(), f"mask{k - 1} not within mask{k}"
    assert l0.min() >= 0 and l0.max() <= L - 1
    print(f"full DMR run: {s0} steps, |mask0| = {int(act.sum())}", flush=True)
