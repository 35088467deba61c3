"""Batched 1-D form (512 seeded shock tubes, uneven sizes) on the GPU.

* The whole batch runs to t = 0.15 in one graph-replayed persistent launch with
  the autotuned shape; representative instances (the largest and smallest
  grids and both thresholds) are bit-identical to independent oracle solves.
* Batched, sequential (one launch per instance) and autotuned runs give
  bit-identical states: batching and tuning change the schedule, not the
  arithmetic.
* One instance of the seeded set (index 448: rho ~ 0.17 on both sides,
  p_L/p_R ~ 0.13, 2049 points, eps 1e-3) breaks down in the reference scheme
  itself: after 1558 steps the next PrepareEuler leaves a non-finite wave
  speed (the oracle, like the reference, carries on until its time turns NaN
  at step 1657).  The GPU is bit-identical to the oracle up to step 1558 and
  then stops that instance with status AWCM_ENONFINITE; every other instance
  reaches t = 0.15.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

BROKEN = 448  # the instance whose state breaks down in the reference scheme


def _modules():
    import pyawcm
    import workloads
    pyawcm.lib()
    return pyawcm, workloads


def oracle_solve(params, q, pts):
    L, nxc, g, dx, tol, cfl, tf = params
    o = O.Solver1D("orc", L, nxc, g, dx)
    o.set_q(q)
    n = 0
    while o.t[0] < tf:
        o.step(tol, cfl, pts)
        n += 1
    return o, n


def test_c4_full_batch_matches_oracle():
    P, W = _modules()
    b, inst, params, states = W.shock_tube_batch(512)
    threads, wgs, probe = b.autotune(10)
    print("autotuned launch", threads, wgs, "probe ms", np.round(probe, 2).tolist())
    assert threads in (256, 512, 1024) and np.all(probe > 0)
    b.run()
    sizes = np.array([p[0] for p in params])
    picks = [int(np.argmax(sizes)), int(np.argmin(sizes))]
    picks += [k for k in range(512) if params[k][4] == 1e-3][:1] + [k for k in range(512) if params[k][4] == 1e-4][:1]
    pts = P.print_times(W.C4_T_FINAL, W.C4_NUM_PRINT)
    for k in picks:
        st = b.state(k)
        o, n = oracle_solve(params[k], states[k], pts)
        assert st["status"] == 0 and st["steps"] == n and st["t"] == o.t[0] == W.C4_T_FINAL, k
        mg, lg = b.masks(k)
        mo, lo = o.masks()
        assert np.array_equal(mg, mo) and np.array_equal(lg, lo), k
        assert np.array_equal(b.download_q(k).view(np.uint64), o.get_q().view(np.uint64)), k
    for k in range(512):
        st = b.state(k)
        if k == BROKEN:
            assert st["status"] == -4 and st["steps"] == 1558 and st["t"] < W.C4_T_FINAL
        else:
            assert st["status"] == 0 and st["t"] == W.C4_T_FINAL, k


def test_batched_sequential_and_tuned_runs_are_identical():
    P, W = _modules()
    results = []
    for mode in ("batched", "sequential", "tuned"):
        b, inst, params, states = W.shock_tube_batch(24, seed=W.C4_SEED + 1)
        if mode == "tuned":
            b.autotune(5)
        if mode == "sequential":
            b.run_sequential(300)
        else:
            b.run(300)
        results.append([b.download_q(k) for k in range(24)] + [np.array([b.state(k)["steps"]]) for k in range(24)])
    for r in results[1:]:
        for x, y in zip(results[0], r):
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64))


def test_broken_instance_matches_oracle_until_breakdown():
    P, W = _modules()
    inst = W.shock_tube_instances(512)
    L, rl, rr, pl, pr, u, x0, eps = inst[BROKEN]
    nx = 16 * 2 ** (L - 1) + 1
    q, dx, _ = P.ic_riemann_1d(nx, 0.0, 1.0, x0, (rl, u, pl), (rr, u, pr))
    params = [(L, 17, 1.4, dx, eps, W.C4_CFL, W.C4_T_FINAL)]
    pts = P.print_times(W.C4_T_FINAL, W.C4_NUM_PRINT)
    b = P.Batch1D(params, pts)
    b.upload_q(0, q)
    b.run(1558)
    o = O.Solver1D("orc", L, 17, 1.4, dx)
    o.set_q(q)
    for _ in range(1558):
        o.step(eps, 0.5, pts)
    st = b.state(0)
    assert st["status"] == 0 and st["steps"] == 1558 and st["t"] == o.t[0]
    assert np.array_equal(b.download_q(0).view(np.uint64), o.get_q().view(np.uint64))
    b.run()
    st = b.state(0)
    assert st["status"] == -4 and st["steps"] == 1558


def test_batch_run_in_pieces_is_one_run():
    """A batch advanced by several calls (graph replays of different lengths,
    with an autotune between them) ends bit-identical to one uninterrupted run
    of the same steps: state carried between calls never changes results."""
    P, W = _modules()
    one, _, _, _ = W.shock_tube_batch(24, seed=W.C4_SEED + 2)
    one.run(300)
    parts, _, _, _ = W.shock_tube_batch(24, seed=W.C4_SEED + 2)
    parts.run(70)
    parts.autotune(3)
    parts.run(130)
    parts.run(100)
    for k in range(24):
        a, b = one.state(k), parts.state(k)
        assert a["steps"] == b["steps"] and a["t"] == b["t"] and a["status"] == b["status"], k
        assert np.array_equal(one.download_q(k).view(np.uint64), parts.download_q(k).view(np.uint64)), k
        ma, la = one.masks(k)
        mb, lb = parts.masks(k)
        assert np.array_equal(ma, mb) and np.array_equal(la, lb), k
