"""The scalar path's time loop on the GPU: awcm_run_scalar runs the box-advection
driver (Wavelet/main_2d_advect.cpp:46-199) -- Prepare, periodic ghosts, three
SSP-RK3 stages with WENO5 right reconstructions and AWCM central differences,
combine, PostProcess -- without returning to the host between iterations.

Checked against
* the reference build's own run of the driver (tests/golden/api_probe_advect.txt:
  hashes of the ghosted field, the five masks and level0..2 after every
  iteration; 40 iterations at 65 x 65, 10 at the driver's 449 x 449);
* the oracle, bit for bit, over longer runs and with the stage fields carried
  across calls.
"""
import numpy as np
import pytest

import advect_lib as A
import oracle_lib as O

pytestmark = pytest.mark.gpu


def _P():
    import pyawcm
    pyawcm.lib()
    return pyawcm


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def device_and_oracle(P, name):
    L, nxc, eps, steps, res, q, dx, dt = A.setup(name)
    qd, dxd = P.ic_box_advection_2d(res)
    assert np.array_equal(bits(qd), bits(q)) and dxd == dx
    S = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nxc, x0=0.0, bc_kind=P.BC_TRANSMISSIVE)
    S.upload_scalar(q)
    xl, xu, _ = O.box_grid(res)
    R = O.Solver2D("orc", L, nxc, nxc, 1.4, 0, xl, xu, xl, xu)
    R.set_scalar(q)
    return S, R, eps, steps, dt


@pytest.mark.parametrize("name", ["small", "driver"])
def test_advect_iterations_match_reference_driver(name):
    P = _P()
    gold = A.golden()
    S, _, eps, steps, dt = device_and_oracle(P, name)
    for s in range(steps):
        S.run_scalar(eps, dt, 1)
        m, lv = S.masks()
        got = A.report(S.download_scalar(), m, lv)
        print(f"{name} iteration {s}: {got}", flush=True)
        assert got == gold[f"{name}.{s}.step"], s


@pytest.mark.parametrize("name,n_calls,per_call", [("small", 4, 25), ("driver", 2, 20)])
def test_advect_matches_oracle_across_calls(name, n_calls, per_call):
    """Several calls of many iterations each: the stage fields carry over
    between calls as the driver's do; field (ghosts included), masks and
    levels bit-identical to the oracle after every call."""
    P = _P()
    S, R, eps, _, dt = device_and_oracle(P, name)
    for c in range(n_calls):
        S.run_scalar(eps, dt, per_call)
        R.advect(eps, dt, per_call)
        m, lv = S.masks()
        mo, lvo = R.masks()
        assert np.array_equal(m, mo) and np.array_equal(lv[:3], lvo[:3]), c
        assert np.array_equal(bits(S.download_scalar()), bits(R.get_scalar())), c
        print(f"{name} call {c}: {(c + 1) * per_call} iterations, mask0 {int(m[1].sum())}", flush=True)
    assert O.load("orc").orc2_oob(R.h) == 0
