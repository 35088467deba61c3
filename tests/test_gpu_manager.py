"""AWCM_MANAGER_1D / AWCM_MANAGER_2D static members on the GPU, through the C
ABI (awcm_add_adjacent_zone, awcm_check_reconstruction,
awcm_add_point_for_differentiation, awcm_compute_level, awcm_remove_residue),
against the oracle on the same arrays: masks of a run and a seeded random
banded mask (as tests/test_oracle.py pins the oracle's to the reference
build's).  Masks and levels bit-exact; the handle's own state unchanged.
"""
import numpy as np
import pytest

import oracle_lib as O
from test_oracle import mask_cases_1d, mask_cases_2d, run_mask_algebra

pytestmark = pytest.mark.gpu


def _pyawcm():
    import pyawcm
    pyawcm.lib()
    return pyawcm


@pytest.mark.parametrize("kind", ["quad", "dmr"])
def test_manager_2d_statics_match_oracle(kind):
    P = _pyawcm()
    L, nxc, nyc = (4, 9, 9) if kind == "quad" else (4, 17, 5)
    nx, ny = (nxc - 1) * 2 ** (L - 1) + 1, (nyc - 1) * 2 ** (L - 1) + 1
    if kind == "quad":
        q, dx, _ = O.ic2(0, nx, ny, 0.0, 1.0, 0.0, 1.0, 0.5)
        bc, xu, tf, npr = 0, 1.0, 0.25, 9
    else:
        q, dx, _ = O.ic2(1, nx, ny, 0.0, 4.0, 0.0, 1.0, 1.4)
        bc, xu, tf, npr = 1, 4.0, 0.2, 21
    o = O.Solver2D("orc", L, nxc, nyc, 1.4, bc, 0.0, xu, 0.0, 1.0)
    o.set_q(q)
    pts = O.print_times(tf, npr)
    for _ in range(5):
        o.step(1e-3, 0.4, pts)
    g = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nyc, x0=0.0, bc_kind=bc)
    g.set_schedule(pts)
    g.upload_q(q)
    g.run(1e-3, 0.4, tf, 5)
    before = g.masks(), g.download_q()
    cases, lv0 = mask_cases_2d(o, np.random.default_rng(41))
    run_mask_algebra(g, o, cases, lv0, 4, o.get_q())
    after = g.masks(), g.download_q()
    assert np.array_equal(before[0][0], after[0][0]) and np.array_equal(before[0][1], after[0][1])
    assert np.array_equal(before[1].view(np.uint64), after[1].view(np.uint64))


@pytest.mark.parametrize("levels", [7, 10])
def test_manager_1d_statics_match_oracle(levels):
    """7 levels (masks in LDS) and 10 levels (8193 points)."""
    P = _pyawcm()
    nx = 16 * 2 ** (levels - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    pts = O.print_times(0.2, 41)
    o = O.Solver1D("orc", levels, 17, 1.4, dx)
    o.set_q(q)
    for _ in range(20):
        o.step(1e-4, 0.5, pts)
    g = P.Solver(1, levels, 17, 1.4, dx)
    g.set_schedule(pts)
    g.upload_q(q)
    g.run(1e-4, 0.5, 0.2, 20)
    before = g.masks(), g.download_q()
    cases, lv0 = mask_cases_1d(o, np.random.default_rng(42), levels)
    run_mask_algebra(g, o, cases, lv0, 3, o.get_q())
    after = g.masks(), g.download_q()
    assert np.array_equal(before[0][0], after[0][0]) and np.array_equal(before[0][1], after[0][1])
    assert np.array_equal(before[1].view(np.uint64), after[1].view(np.uint64))


def test_manager_rejects_levels_out_of_range():
    P = _pyawcm()
    g = P.Solver(1, 5, 17, 1.4, 1.0 / 256)
    m = np.ones(257 + 6, np.uint8)
    with pytest.raises(P.AwcmError) as e:
        g.mask_op(2, m, np.full(257, 7))
    assert e.value.code == -1
