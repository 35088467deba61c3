"""GPU parity: the HIP solver (through the C ABI) against the CPU oracle.

Bar: masks and level fields bit-exact, fields bit-exact (the FP64 operation
order is the reference's and no FMA contraction is allowed), dt and t equal.
The per-field relative-L1 figure is also asserted against 1e-12 so a failure
message shows how far off a field is.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _pyawcm():
    import pyawcm
    pyawcm.lib()  # raises if the HIP library is not built
    return pyawcm


def rel_l1(a, b):
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / (den if den > 0 else 1.0))


def assert_same(qg, qo, mg, mo, lg, lo, what):
    assert np.array_equal(mg, mo), f"{what}: masks differ at {np.argwhere(mg != mo)[:5].tolist()}"
    assert np.array_equal(lg, lo), f"{what}: levels differ at {np.argwhere(lg != lo)[:5].tolist()}"
    assert rel_l1(qg, qo) <= 1e-12, f"{what}: rel L1 {rel_l1(qg, qo)}"
    assert np.array_equal(qg.view(np.uint64), qo.view(np.uint64)), \
        f"{what}: fields not bit-identical (max abs diff {np.abs(qg - qo).max()})"


def sod_setup(L=7, nxc=17):
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    return q, dx


def test_1d_prepare_matches_oracle():
    P = _pyawcm()
    q, dx = sod_setup()
    g = P.Solver(1, 7, 17, 1.4, dx)
    o = O.Solver1D("orc", 7, 17, 1.4, dx)
    g.upload_q(q)
    o.set_q(q)
    g.prepare(1e-4)
    o.prepare(1e-4)
    mg, lg = g.masks()
    mo, lo = o.masks()
    assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, "1-D prepare")


def test_1d_sod_steps_match_oracle():
    P = _pyawcm()
    q, dx = sod_setup()
    pts = O.print_times(0.2, 41)
    g = P.Solver(1, 7, 17, 1.4, dx)
    g.set_schedule(pts)
    o = O.Solver1D("orc", 7, 17, 1.4, dx)
    g.upload_q(q)
    o.set_q(q)
    for n in range(12):
        o.step(1e-4, 0.5, pts)
        g.run(1e-4, 0.5, 0.2, 1)
        assert g.t == o.t[0] and g.cnt == o.cnt.value, f"step {n}: t {g.t} vs {o.t[0]}"
        mg, lg = g.masks()
        mo, lo = o.masks()
        assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, f"1-D step {n}")


def test_1d_sod_full_run_matches_oracle():
    P = _pyawcm()
    q, dx = sod_setup()
    pts = O.print_times(0.2, 41)
    g = P.Solver(1, 7, 17, 1.4, dx)
    g.set_schedule(pts)
    g.upload_q(q)
    steps, upd = g.run(1e-4, 0.5, 0.2, 100000)
    o = O.Solver1D("orc", 7, 17, 1.4, dx)
    o.set_q(q)
    n = 0
    while o.t[0] < 0.2:
        o.step(1e-4, 0.5, pts)
        n += 1
    assert steps == n and g.t == o.t[0] == 0.2
    assert upd > 0
    mg, lg = g.masks()
    mo, lo = o.masks()
    assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, "1-D Sod t=0.2")


def test_1d_weno5_flux_matches_oracle():
    """The shipped driver's stage flux (WENO_EULER_SOLVER_1D::ComputedF,
    main_1d_euler.cpp:109) on the device, awcm_set_flux(AWCM_FLUX_WENO5): a
    full Sod run (7 levels, 1025 points) and 200 steps in the driver's own
    configuration (SHOCKTUBE on [0, 10], 11 levels, 16 coarse points, 15361
    points: masks in HBM).  Masks, levels, step counts, t and fields
    bit-identical to the oracle's WENO5-JS run.  (Step 406 of the Sod run
    divides E + p = 0x1.6666666666663p-2 by rho = 0x1.ffffffffffffbp-4, whose
    exact quotient lies 0.5000000000000001 ulp above 0x1.6666666666666p+1: the
    hardware divide sequence returns 0x1.6666666666666p+1, the 1-D path's
    correctly rounded quotients (awcm::Div) return 0x1.6666666666667p+1, as
    the oracle does.)  The adjoint differentiates the WENO-Z pair and
    refuses a WENO5-JS handle."""
    P = _pyawcm()
    q, dx = sod_setup()
    pts = O.print_times(0.2, 41)
    g = P.Solver(1, 7, 17, 1.4, dx)
    g.set_flux(P.FLUX_WENO5)
    g.set_schedule(pts)
    g.upload_q(q)
    steps, _ = g.run(1e-4, 0.5, 0.2, 100000)
    o = O.Solver1D("orc", 7, 17, 1.4, dx)
    o.set_flux(1)
    o.set_q(q)
    n = 0
    while o.t[0] < 0.2:
        o.step(1e-4, 0.5, pts)
        n += 1
    assert steps == n and g.t == o.t[0] == 0.2
    mg, lg = g.masks()
    mo, lo = o.masks()
    assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, "1-D Sod WENO5-JS t=0.2")
    with pytest.raises(P.AwcmError) as e:
        g.adjoint_record(1e-4, 0.5, 0.3, 4)
    assert e.value.code == -6  # AWCM_ESTATE
    # main_1d_euler.cpp's configuration (:14-17)
    L, nxc = 11, 16
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 10.0, 5.0, O.SOD_L, O.SOD_R)
    pts = O.print_times(2.0, 41)
    g = P.Solver(1, L, nxc, 1.4, dx)
    g.set_flux(P.FLUX_WENO5)
    g.set_schedule(pts)
    g.upload_q(q)
    o = O.Solver1D("orc", L, nxc, 1.4, dx)
    o.set_flux(1)
    o.set_q(q)
    for chunk in range(4):
        steps, _ = g.run(1e-6, 0.5, 2.0, 50)
        for _ in range(50):
            o.step(1e-6, 0.5, pts)
        assert steps == 50 and g.t == o.t[0], f"chunk {chunk}"
        mg, lg = g.masks()
        mo, lo = o.masks()
        assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, f"driver configuration, {50 * (chunk + 1)} steps")


def test_1d_state_in_lds_matches_oracle():
    """Single instances too long for their whole state in one workgroup's LDS
    but whose masks and Q fit (2049 and 4097 points): Q lives in LDS for the
    run.  The full Shu-Osher run (9 levels, 4097 points, eps 1e-4, CFL 0.5,
    t = 1.8, WENO-Z pair, checked at t = 0.9 and 1.8) and a full Sod run at
    8 levels (2049 points, WENO5-JS): masks, levels, step counts, t and
    fields bit-identical to the oracle."""
    P = _pyawcm()
    q, dx, _ = O.ic1(1, 4097, -5.0, 5.0, -4.0, O.SHU_L, O.SHU_L)
    pts = O.print_times(1.8, 3)
    g = P.Solver(1, 9, 17, 1.4, dx)
    g.set_schedule(pts)
    g.upload_q(q)
    o = O.Solver1D("orc", 9, 17, 1.4, dx)
    o.set_q(q)
    for tf in (0.9, 1.8):
        steps, _ = g.run(1e-4, 0.5, tf, 1 << 20)
        n = 0
        while o.t[0] < tf:
            o.step(1e-4, 0.5, pts)
            n += 1
        assert steps == n and g.t == o.t[0] == tf
        mg, lg = g.masks()
        mo, lo = o.masks()
        assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, f"Shu-Osher 4097 points t={tf}")
    q, dx, _ = O.ic1(0, 2049, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    pts = O.print_times(0.2, 41)
    g = P.Solver(1, 8, 17, 1.4, dx)
    g.set_flux(P.FLUX_WENO5)
    g.set_schedule(pts)
    g.upload_q(q)
    steps, _ = g.run(1e-4, 0.5, 0.2, 100000)
    o = O.Solver1D("orc", 8, 17, 1.4, dx)
    o.set_flux(1)
    o.set_q(q)
    n = 0
    while o.t[0] < 0.2:
        o.step(1e-4, 0.5, pts)
        n += 1
    assert steps == n and g.t == o.t[0] == 0.2
    mg, lg = g.masks()
    mo, lo = o.masks()
    assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, "Sod 2049 points WENO5-JS t=0.2")


def test_1d_largest_instance_matches_oracle():
    """The largest 1-D instance the C ABI accepts with the most levels: 14
    levels over 8 coarse points, 57345 finest points (list entries pack the
    index in 16 bits, so a grid stops at 65535 points; masks in HBM).  A Sod
    tube on [0, 1], three runs of 10 steps, masks, levels, t and fields
    bit-identical to the oracle after each."""
    P = _pyawcm()
    L, nxc = 14, 8
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    assert nx == 57345
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    pts = O.print_times(0.2, 41)
    g = P.Solver(1, L, nxc, 1.4, dx)
    g.set_schedule(pts)
    g.upload_q(q)
    o = O.Solver1D("orc", L, nxc, 1.4, dx)
    o.set_q(q)
    for chunk in range(3):
        steps, _ = g.run(1e-4, 0.5, 0.2, 10)
        for _ in range(10):
            o.step(1e-4, 0.5, pts)
        assert steps == 10 and g.t == o.t[0], f"chunk {chunk}"
        mg, lg = g.masks()
        mo, lo = o.masks()
        assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, f"57345 points, {10 * (chunk + 1)} steps")


def test_batch_weno5_flux_matches_oracle():
    """awcm_batch_set_flux(AWCM_FLUX_WENO5) on a batch of six Riemann
    problems; then back to WENO-Z for a further run (the batch graph is
    re-captured for the other kernel).  Bit-identical to the oracle."""
    P = _pyawcm()
    rng = np.random.default_rng(9)
    params, qs = [], []
    for k in range(6):
        L = 6 + k % 2
        nx = 16 * 2 ** (L - 1) + 1
        rl, rr = rng.uniform(0.1, 1.0, 2)
        pl, pr = rng.uniform(0.05, 1.0, 2)
        q, dx, _ = O.ic1(0, nx, 0.0, 1.0, rng.uniform(0.3, 0.7), (rl, 0.0, pl), (rr, 0.0, pr))
        params.append((L, 17, 1.4, dx, 1e-4, 0.5, 0.04))
        qs.append(q)
    pts = O.print_times(0.04, 5)
    b = P.Batch1D(params, pts)
    for k, q in enumerate(qs):
        b.upload_q(k, q)
    b.set_flux(P.FLUX_WENO5)
    b.run(30)
    b.set_flux(P.FLUX_WENOZ)
    b.run(20)
    for k, q in enumerate(qs):
        L, nxc, gm, dx, tol, cfl, tf = params[k]
        o = O.Solver1D("orc", L, nxc, gm, dx)
        o.set_q(q)
        o.set_flux(1)
        for _ in range(30):
            if o.t[0] < tf:
                o.step(tol, cfl, pts)
        o.set_flux(0)
        for _ in range(20):
            if o.t[0] < tf:
                o.step(tol, cfl, pts)
        st = b.state(k)
        assert st["status"] == 0 and st["t"] == o.t[0], k
        mg, lg = b.masks(k)
        mo, lo = o.masks()
        assert_same(b.download_q(k), o.get_q(), mg, mo, lg, lo, f"batch instance {k}")


def quad_setup(L, nxc):
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    q, dx, _ = O.ic2(0, nx, nx, 0.0, 1.0, 0.0, 1.0, 0.5)
    return q, dx


def dmr_setup(L, nxc, nyc, rho_pre=1.4):
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    ny = (nyc - 1) * 2 ** (L - 1) + 1
    q, dx, _ = O.ic2(1, nx, ny, 0.0, 4.0, 0.0, 1.0, rho_pre)
    return q, dx


@pytest.mark.parametrize("case", ["quad", "dmr"])
def test_2d_prepare_matches_oracle(case):
    P = _pyawcm()
    if case == "quad":
        L, nxc, nyc, bc = 5, 9, 9, 0
        q, dx = quad_setup(L, nxc)
        xl, xu = 0.0, 1.0
    else:
        L, nxc, nyc, bc = 4, 33, 9, 1
        q, dx = dmr_setup(L, nxc, nyc)
        xl, xu = 0.0, 4.0
    g = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nyc, x0=xl, bc_kind=bc)
    o = O.Solver2D("orc", L, nxc, nyc, 1.4, bc, xl, xu, 0.0, 1.0)
    g.upload_q(q)
    o.set_q(q)
    g.prepare(1e-3)
    o.prepare(1e-3)
    mg, lg = g.masks()
    mo, lo = o.masks()
    assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, f"2-D {case} prepare")


@pytest.mark.parametrize("tail_cells", [4096, 0])
@pytest.mark.parametrize("case", ["quad", "dmr"])
def test_2d_steps_match_oracle(case, tail_cells):
    P = _pyawcm()
    if case == "quad":
        L, nxc, nyc, bc, tf, npr = 5, 9, 9, 0, 0.25, 9
        q, dx = quad_setup(L, nxc)
        xl, xu = 0.0, 1.0
    else:
        L, nxc, nyc, bc, tf, npr = 4, 33, 9, 1, 0.2, 21
        q, dx = dmr_setup(L, nxc, nyc)
        xl, xu = 0.0, 4.0
    pts = O.print_times(tf, npr)
    g = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nyc, x0=xl, bc_kind=bc)
    g.set_tuning(P.TUNE_TAIL_CELLS, tail_cells)
    g.set_schedule(pts)
    o = O.Solver2D("orc", L, nxc, nyc, 1.4, bc, xl, xu, 0.0, 1.0)
    g.upload_q(q)
    o.set_q(q)
    for n in range(15):
        o.step(1e-3, 0.4, pts)
        g.run(1e-3, 0.4, tf, 1)
        assert g.t == o.t[0] and g.cnt == o.cnt.value, f"step {n}: t {g.t} vs {o.t[0]}"
        mg, lg = g.masks()
        mo, lo = o.masks()
        assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, f"2-D {case} step {n}")


@pytest.mark.parametrize("shape", [(4096, 5), (5, 4096)])
def test_2d_index_limit_strip_matches_oracle(shape):
    """Rows and columns at the 16-bit limit of the packed list entries: a
    65521 x 65 strip and its transpose (5 levels), square cells, the quadrant
    problem with transmissive boundaries.  Two steps, masks, levels, t and
    fields bit-identical to the oracle after each (on the CPU the oracle
    matches the reference build on the same two steps, both ways round)."""
    P = _pyawcm()
    L, (nxc, nyc) = 5, shape
    nx, ny = (nxc - 1) * 16 + 1, (nyc - 1) * 16 + 1
    assert max(nx, ny) == 65521
    h = 1.0 / (max(nx, ny) - 1)
    xu, yu = (nx - 1) * h, (ny - 1) * h
    q, dx, _ = O.ic2(0, nx, ny, 0.0, xu, 0.0, yu, 0.5)
    pts = O.print_times(0.25, 9)
    g = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nyc, x0=0.0, bc_kind=0)
    g.set_schedule(pts)
    o = O.Solver2D("orc", L, nxc, nyc, 1.4, 0, 0.0, xu, 0.0, yu)
    g.upload_q(q)
    o.set_q(q)
    for n in range(2):
        o.step(1e-3, 0.4, pts)
        g.run(1e-3, 0.4, 0.25, 1)
        assert g.t == o.t[0] and g.cnt == o.cnt.value, f"step {n}"
        mg, lg = g.masks()
        mo, lo = o.masks()
        assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, f"{nx} x {ny} strip, step {n}")


@pytest.mark.parametrize("tail_cells", [4096, 256, 0])
def test_2d_multistep_run_matches_oracle(tail_cells):
    """Several steps queued in one awcm_run call (graph replays, no host round
    trip per step), with every level cascade split between per-level kernels
    and the one-workgroup tail in three different ways."""
    P = _pyawcm()
    L, nxc, nyc = 5, 33, 9
    q, dx = dmr_setup(L, nxc, nyc)
    pts = O.print_times(0.2, 21)
    g = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nyc, x0=0.0, bc_kind=1)
    g.set_tuning(P.TUNE_TAIL_CELLS, tail_cells)
    g.set_schedule(pts)
    g.upload_q(q)
    steps, _ = g.run(1e-3, 0.4, 0.2, 40)
    o = O.Solver2D("orc", L, nxc, nyc, 1.4, 1, 0.0, 4.0, 0.0, 1.0)
    o.set_q(q)
    for _ in range(steps):
        o.step(1e-3, 0.4, pts)
    assert steps == 40 and g.t == o.t[0]
    mg, lg = g.masks()
    mo, lo = o.masks()
    assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, "2-D DMR 40 steps")


def test_2d_run_past_t_final_keeps_the_last_steps_masks():
    """A run whose step budget outlasts t_final queues steps that do nothing
    (t >= t_final): they must leave every mask as the last real step left it
    -- mask_temp (mask0 before that step's PrepareEuler) and masks 1..3 --
    not only the field, mask0 and the levels."""
    P = _pyawcm()
    L, nxc, nyc = 5, 33, 9
    q, dx = dmr_setup(L, nxc, nyc)
    pts = O.print_times(0.2, 21)
    g = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nyc, x0=0.0, bc_kind=1)
    g.set_schedule(pts)
    g.upload_q(q)
    o = O.Solver2D("orc", L, nxc, nyc, 1.4, 1, 0.0, 4.0, 0.0, 1.0)
    o.set_q(q)
    for k in (1, 2):
        steps, _ = g.run(1e-3, 0.4, pts[k], 4096)
        n = 0
        while o.t[0] < pts[k]:
            o.step(1e-3, 0.4, pts)
            n += 1
        assert steps == n and g.t == o.t[0] and g.cnt == o.cnt.value
        mg, lg = g.masks()
        mo, lo = o.masks()
        assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, f"2-D DMR run to output time {k} ({n} steps)")


def test_2d_prepare_then_run_matches_oracle():
    """awcm_prepare_euler on its own, then steps: the standalone PrepareEuler
    drops points from mask0 without a residue pass, so the next step's
    residue pass must zero every point outside mask0 (as RemoveResidueEuler
    always does), not only the points the step's masks touch."""
    P = _pyawcm()
    L, nxc, nyc = 5, 33, 9
    q, dx = dmr_setup(L, nxc, nyc)
    pts = O.print_times(0.2, 21)
    g = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nyc, x0=0.0, bc_kind=1)
    g.set_schedule(pts)
    o = O.Solver2D("orc", L, nxc, nyc, 1.4, 1, 0.0, 4.0, 0.0, 1.0)
    g.upload_q(q)
    o.set_q(q)
    for rep in range(2):
        g.prepare(1e-3)
        o.prepare(1e-3)
        steps, _ = g.run(1e-3, 0.4, 0.2, 5)
        for _ in range(5):
            o.step(1e-3, 0.4, pts)
        assert steps == 5 and g.t == o.t[0]
        mg, lg = g.masks()
        mo, lo = o.masks()
        assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, f"2-D DMR prepare then 5 steps, round {rep}")


@pytest.mark.parametrize("case", ["dmr", "quad"])
def test_2d_full_size_matches_oracle(case):
    """The 2-D headline configurations at their full sizes: double Mach
    reflection on 4097 x 1025 (7 levels, 65 x 17 coarse, rho_pre 1.4) and the
    four-quadrant Riemann problem on 2049 x 2049 (7 levels, 33 x 33 coarse),
    eps 1e-3, CFL 0.4.  PrepareEuler of the initial condition on its own; then,
    from the initial condition again as the drivers run it, 30 steps queued 10
    at a time (graph replays).  Masks, levels, t, the print counter and all four
    fields are bit-identical to the oracle after the PrepareEuler and after
    every 10 steps.  (A PrepareEuler on its own followed by steps applies the
    transform twice before the first stage; on the four-quadrant state the
    reference scheme breaks down at once that way, and so do oracle and GPU.)"""
    P = _pyawcm()
    if case == "dmr":
        L, nxc, nyc, bc, xl, xu, tf, npr = 7, 65, 17, 1, 0.0, 4.0, 0.2, 21
        q, dx = dmr_setup(L, nxc, nyc)
    else:
        L, nxc, nyc, bc, xl, xu, tf, npr = 7, 33, 33, 0, 0.0, 1.0, 0.25, 9
        q, dx = quad_setup(L, nxc)
    pts = O.print_times(tf, npr)

    def pair():
        g = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nyc, x0=xl, bc_kind=bc)
        g.set_schedule(pts)
        o = O.Solver2D("orc", L, nxc, nyc, 1.4, bc, xl, xu, 0.0, 1.0)
        g.upload_q(q)
        o.set_q(q)
        return g, o

    g, o = pair()
    g.prepare(1e-3)
    o.prepare(1e-3)
    mg, lg = g.masks()
    mo, lo = o.masks()
    assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, f"2-D {case} full size prepare")
    del g, o
    g, o = pair()
    for chunk in range(3):
        steps, _ = g.run(1e-3, 0.4, tf, 10)
        for _ in range(10):
            o.step(1e-3, 0.4, pts)
        assert steps == 10 and g.t == o.t[0] and g.cnt == o.cnt.value, f"chunk {chunk}: t {g.t} vs {o.t[0]}"
        mg, lg = g.masks()
        mo, lo = o.masks()
        qg = g.download_q()
        assert np.isfinite(qg).all()
        assert_same(qg, o.get_q(), mg, mo, lg, lo, f"2-D {case} full size, {10 * (chunk + 1)} steps")


@pytest.mark.parametrize("case", ["quad", "dmr", "dmr_full"])
def test_2d_interpolate_all_grid_matches_oracle(case):
    """WAVELET_2D::InterpolateAllGridEuler through awcm_interpolate_all_grid.
    After a run (GPU and oracle in step; the last case is double Mach
    reflection at 4097 x 1025), both interpolate the same field from mask0
    (the driver's call, main_2d_euler.cpp:334), mask1, mask3, a sparse and a
    dense random mask and no points at all: fields bit-identical.  Then, from
    the driver's call, the run goes on on both (a field that is non-zero
    outside mask0 entering PrepareEuler): masks, levels, fields, t
    bit-identical, for as long as the reference scheme keeps a physical
    state (see below)."""
    P = _pyawcm()
    if case == "quad":
        L, nxc, nyc, bc, xl, xu, tf, npr, n0 = 5, 9, 9, 0, 0.0, 1.0, 0.25, 9, 8
        q, dx = quad_setup(L, nxc)
    elif case == "dmr":
        L, nxc, nyc, bc, xl, xu, tf, npr, n0 = 4, 33, 9, 1, 0.0, 4.0, 0.2, 21, 8
        q, dx = dmr_setup(L, nxc, nyc)
    else:
        L, nxc, nyc, bc, xl, xu, tf, npr, n0 = 7, 65, 17, 1, 0.0, 4.0, 0.2, 21, 10
        q, dx = dmr_setup(L, nxc, nyc)
    pts = O.print_times(tf, npr)
    rng = np.random.default_rng(17)
    g = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nyc, x0=xl, bc_kind=bc)
    g.set_schedule(pts)
    g.upload_q(q)
    o = O.Solver2D("orc", L, nxc, nyc, 1.4, bc, xl, xu, 0.0, 1.0)
    steps, _ = g.run(1e-3, 0.4, tf, n0)
    assert steps == n0
    o.set_q(q)
    for _ in range(n0):
        o.step(1e-3, 0.4, pts)
    assert g.t == o.t[0]
    mg, lg = g.masks()
    mo, lo = o.masks()
    qs = g.download_q()
    assert_same(qs, o.get_q(), mg, mo, lg, lo, f"{case} after {n0} steps")
    froms = {"mask1": mg[2], "mask3": mg[4], "sparse": (rng.random(mg[1].shape) < 0.1).astype(np.uint8),
             "dense": (rng.random(mg[1].shape) < 0.7).astype(np.uint8), "none": np.zeros_like(mg[1]),
             "mask0": mg[1]}
    for name, f in froms.items():
        print(f"all grid {case} from {name}", flush=True)
        g.upload_q(qs)
        o.set_q(qs)
        g.interpolate_all_grid(f)
        o.interpolate_all(f)
        qg, qo = g.download_q(), o.get_q()
        assert np.array_equal(qg.view(np.uint64), qo.view(np.uint64)), \
            f"{case} from {name}: {int((qg != qo).sum())} values differ"
    # the field of the driver's call (from mask0) goes on.  The interpolated
    # points feed the next transforms, and on two of the cases the reference
    # scheme then leaves the physical states: on the four-quadrant state the
    # second step gives NaN (the reference build too, at the same points as
    # the oracle), and on the full-size double Mach state the second step's
    # PrepareEuler activates points of negative pressure, whose NaN sound
    # speed makes the reference's maximum of the wave speeds (so dt) depend on
    # the order its loop visits them.  Up to there the fields are bit-identical;
    # on the four-quadrant case the GPU field must hold NaN on the same step.
    for n in range(1 if case == "dmr_full" else 3):
        o.step(1e-3, 0.4, pts)
        g.run(1e-3, 0.4, tf, 1)
        what = f"{case} step {n} after the interpolation"
        assert g.t == o.t[0] and g.cnt == o.cnt.value, what
        mg2, lg2 = g.masks()
        mo2, lo2 = o.masks()
        qg, qo = g.download_q(), o.get_q()
        nan = np.isnan(qo)
        if nan.any():
            assert np.isnan(qg).any(), f"{what}: the oracle broke down, the GPU did not"
            break
        assert_same(qg, qo, mg2, mo2, lg2, lo2, what)


def test_set_flux_rejects_2d_handle():
    """The stage-flux switch is the 1-D solver's (the 2-D drivers use WENO-Z):
    a 2-D handle is refused with AWCM_ESTATE."""
    P = _pyawcm()
    q, dx = dmr_setup(4, 33, 9)
    g = P.Solver(2, 4, 33, 1.4, dx, ny_coarse=9, x0=0.0, bc_kind=1)
    with pytest.raises(P.AwcmError) as e:
        g.set_flux(P.FLUX_WENO5)
    assert e.value.code == -6  # AWCM_ESTATE


def test_1d_interpolate_euler_matches_oracle():
    """WAVELET_1D::InterpolateEuler through the C ABI: the two-mask form
    (awcm_interpolate_euler: momentum and energy) and the one-mask form
    (awcm_interpolate_all_grid on a 1-D handle: every point, all three
    components), on a Sod state after 20 steps; fields bit-identical to the
    oracle and the instance's masks unchanged."""
    P = _pyawcm()
    q, dx = sod_setup()
    pts = O.print_times(0.2, 41)
    g = P.Solver(1, 7, 17, 1.4, dx)
    g.set_schedule(pts)
    g.upload_q(q)
    g.run(1e-4, 0.5, 0.2, 20)
    o = O.Solver1D("orc", 7, 17, 1.4, dx)
    o.set_q(q)
    for _ in range(20):
        o.step(1e-4, 0.5, pts)
    qs = g.download_q()
    mg, lg = g.masks()
    mo, lo = o.masks()
    assert_same(qs, o.get_q(), mg, mo, lg, lo, "1-D after 20 steps")
    rng = np.random.default_rng(6)
    m = mg
    for name, (f, t) in {"m0_all": (m[1], None), "m0_m3": (m[1], m[4]), "m1_m2": (m[2], m[3]),
                         "random_all": ((rng.random(m[1].shape) < 0.3).astype(np.uint8), None)}.items():
        g.upload_q(qs)
        o.set_q(qs)
        if t is None:
            g.interpolate_all_grid(f)
        else:
            g.interpolate_euler(f, t)
        o.interpolate_euler(f, t)
        a, b = g.download_q(), o.get_q()
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), f"InterpolateEuler {name}: {int((a != b).sum())} differ"
        mg2, lg2 = g.masks()
        assert np.array_equal(mg2, mg) and np.array_equal(lg2, lg), f"InterpolateEuler {name} changed the masks"


@pytest.mark.parametrize("L,nxc", [(7, 17), (8, 17), (11, 16)])
def test_1d_scalar_path_matches_oracle(L, nxc):
    """The 1-D scalar path through the C ABI: AWCM_1D::Prepare (awcm_prepare)
    over four rounds with the field moved between them, then
    WAVELET_1D::Interpolate (awcm_interpolate; the one-mask form is `to` =
    every point) from/to five mask pairs; the instance's own masks are the
    same afterwards.  Masks, levels and fields bit-identical to the oracle.
    The grids place the state differently: 1025 points keep all of it in LDS,
    2049 points the field alone, 15361 points keep even the masks in HBM."""
    P = _pyawcm()
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    x = (np.arange(-3, nx + 3) + 0.5) / nx
    q = np.exp(-200 * (x - 0.4) ** 2) + 0.5 * (np.abs(x - 0.7) < 0.1)
    g = P.Solver(1, L, nxc, 1.4, 1.0 / nx)
    o = O.Solver1D("orc", L, nxc, 1.4, 1.0 / nx)
    cur = q
    for r in range(4):
        print(f"scalar 1-D L{L} prepare round {r}", flush=True)
        g.upload_scalar(cur)
        o.set_scalar(cur)
        g.prepare_scalar(1e-4)
        o.prepare_scalar(1e-4)
        mg, lg = g.masks()
        mo, lo = o.masks()
        sg = g.download_scalar()
        assert not mg[0].any(), "the scalar Prepare leaves mask_temp alone"
        assert_same(sg, o.get_scalar(), mg, mo, lg, lo, f"1-D scalar prepare round {r}")
        cur = np.roll(sg, 2) * 0.999
    rng = np.random.default_rng(4)
    m = mg
    ones = np.ones_like(m[1])
    for name, (f, t) in {"m0_all": (m[1], None), "m0_m3": (m[1], m[4]), "m1_m2": (m[2], m[3]),
                         "random": ((rng.random(m[1].shape) < 0.3).astype(np.uint8),
                                    (rng.random(m[1].shape) < 0.6).astype(np.uint8)),
                         "sparse_all": ((rng.random(m[1].shape) < 0.2).astype(np.uint8), None)}.items():
        print(f"scalar 1-D L{L} interpolate {name}", flush=True)
        g.upload_scalar(q)
        o.set_scalar(q)
        g.interpolate(f, ones if t is None else t)
        o.interpolate_scalar(f, t)
        a, b = g.download_scalar(), o.get_scalar()
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), f"Interpolate {name}: {int((a != b).sum())} differ"
        mg2, lg2 = g.masks()
        assert np.array_equal(mg2, mg) and np.array_equal(lg2, lg), f"Interpolate {name} changed the masks"


@pytest.mark.parametrize("case", ["box", "dmr_density", "box_2049"])
def test_2d_scalar_path_matches_oracle(case):
    """The scalar path through the C ABI: AWCM_2D::Prepare (awcm_prepare) on a
    scalar field, four rounds with the field moved between them (shifted one
    column and damped, a stand-in for the driver's stages), then
    WAVELET_2D::Interpolate (awcm_interpolate) and InterpolateEuler
    (awcm_interpolate_euler) from/to five mask pairs.  Masks, levels and
    fields bit-identical to the oracle; the scalar Prepare leaves mask_temp
    alone.  The last case is a 2049 x 2049 grid (7 levels, 33 x 33 coarse)."""
    P = _pyawcm()
    if case == "dmr_density":
        L, nxc, nyc, xu, rounds = 4, 33, 9, 4.0, 4
    elif case == "box":
        L, nxc, nyc, xu, rounds = 5, 9, 9, 1.0, 4
    else:
        L, nxc, nyc, xu, rounds = 7, 33, 33, 1.0, 2
    nx, ny = (nxc - 1) * 2 ** (L - 1) + 1, (nyc - 1) * 2 ** (L - 1) + 1
    dx = xu / (nx - 1)
    if case == "dmr_density":
        q = O.ic2(1, nx, ny, 0.0, 4.0, 0.0, 1.0, 1.4)[0].reshape(-1, 4)[:, 0].copy()
    else:
        x = (np.arange(-3, nx + 3) + 0.5) / nx
        X, Y = np.meshgrid(x, x)
        q = (np.exp(-50 * (X - 0.5) ** 2 - 50 * (Y - 0.5) ** 2) + 0.3 * (X > 0.7)).ravel()
    g = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nyc, x0=0.0, bc_kind=0)
    o = O.Solver2D("orc", L, nxc, nyc, 1.4, 0, 0.0, xu, 0.0, 1.0)
    cur = q
    for r in range(rounds):
        print(f"scalar {case} prepare round {r}", flush=True)
        g.upload_scalar(cur)
        o.set_scalar(cur)
        g.prepare_scalar(1e-4)
        o.prepare_scalar(1e-4)
        mg, lg = g.masks()
        mo, lo = o.masks()
        sg = g.download_scalar()
        assert not mg[0].any(), "the scalar Prepare leaves mask_temp alone"
        assert_same(sg, o.get_scalar(), mg, mo, lg, lo, f"scalar {case} prepare round {r}")
        cur = (np.roll(sg.reshape(ny + 6, nx + 6), 1, axis=1) * 0.999).ravel()
    rng = np.random.default_rng(21)
    m = mg
    qe = np.stack([q, 2 * q, q - 1, 3 * q + 1], axis=-1).ravel()
    pairs = {"m0_all": (m[1], np.ones_like(m[1])), "m0_m3": (m[1], m[4]), "m1_m2": (m[2], m[3]),
             "none_m3": (np.zeros_like(m[1]), m[4]),
             "random": ((rng.random(m[1].shape) < 0.3).astype(np.uint8), (rng.random(m[1].shape) < 0.6).astype(np.uint8))}
    for name, (f, t) in pairs.items():
        print(f"scalar {case} interpolate {name}", flush=True)
        g.upload_scalar(q)
        o.set_scalar(q)
        g.interpolate(f, t)
        o.interpolate_scalar(f, t)
        a, b = g.download_scalar(), o.get_scalar()
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), f"Interpolate {name}: {int((a != b).sum())} differ"
        g.upload_q(qe)
        o.set_q(qe)
        g.interpolate_euler(f, t)
        o.interpolate_euler(f, t)
        a, b = g.download_q(), o.get_q()
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), f"InterpolateEuler {name}: {int((a != b).sum())} differ"


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4, 5])
def test_2d_api_sequences_match_oracle(seed):
    """Seeded random sequences of C-ABI calls on one solver, mirrored on the
    oracle and compared bit for bit after every call: PrepareEuler on its own,
    queued runs of 1..5 steps, uploads of a perturbed state, mask0 pushed back,
    InterpolateAllGridEuler from mask0, and (GPU only, no effect on results)
    the tail and graph-replay settings.
    State carried between calls (the residue flag, mask_temp, captured graphs)
    must never make the result depend on the call sequence.  Where the
    reference scheme breaks down (repeated PrepareEuler before a step can),
    the GPU must report it in the same run of steps as the oracle."""
    P = _pyawcm()
    rng = np.random.default_rng(seed)
    L, nxc, nyc, tf = 4, 33, 9, 0.2
    q, dx = dmr_setup(L, nxc, nyc)
    pts = O.print_times(tf, 21)
    g = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nyc, x0=0.0, bc_kind=1)
    g.set_schedule(pts)
    o = O.Solver2D("orc", L, nxc, nyc, 1.4, 1, 0.0, 4.0, 0.0, 1.0)
    g.upload_q(q)
    o.set_q(q)
    ops = []
    for _ in range(16):
        k = int(rng.integers(0, 7))
        if k == 0:
            ops.append("prepare")
            g.prepare(1e-3)
            o.prepare(1e-3)
        elif k == 6:
            ops.append("all grid from mask0")
            m0 = o.masks()[0][1].copy()
            g.interpolate_all_grid(m0)
            o.interpolate_all(m0)
        elif k in (1, 2):
            n = int(rng.integers(1, 6))
            ops.append(f"run {n}")
            try:
                steps, _ = g.run(1e-3, 0.4, tf, n)
            except P.AwcmError as e:
                # the reference scheme itself breaks down on some sequences
                # (repeated PrepareEuler before a step): so must the oracle
                assert "ENONFINITE" in str(e), ops
                broke = False
                for _ in range(n):
                    dt = o.step(1e-3, 0.4, pts)
                    if not (np.isfinite(dt) and np.isfinite(o.get_q()).all()):
                        broke = True
                        break
                assert broke, f"seed {seed}: GPU broke down after {ops}, the oracle did not"
                return
            for _ in range(steps):
                o.step(1e-3, 0.4, pts)
            assert g.t == o.t[0] and g.cnt == o.cnt.value, ops
        elif k == 3:
            ops.append("upload perturbed")
            qn = g.download_q() * (1.0 + 1e-3 * rng.standard_normal(g.npts * g.ncomp))
            g.upload_q(qn)
            o.set_q(qn)
        elif k == 4:
            ops.append("mask0 pushed back")
            m0 = o.masks()[0][1].copy()
            g.set_mask0(m0)
            o.set_mask0(m0)
        else:
            tail, graph = int(rng.choice([0, 256, 4096])), int(rng.integers(0, 2))
            ops.append(f"tail {tail} graph {graph}")
            g.set_tuning(P.TUNE_TAIL_CELLS, tail)
            g.set_tuning(P.TUNE_GRAPH, graph)
        mg, lg = g.masks()
        mo, lo = o.masks()
        assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, f"seed {seed} after {ops}")


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_1d_api_sequences_match_oracle(seed):
    """The 1-D solver under seeded random sequences of C-ABI calls, mirrored
    on the oracle and compared bit for bit after every call: PrepareEuler on
    its own, queued runs of 1..8 steps, uploads of a perturbed state and mask0
    pushed back.  A breakdown of the reference scheme must be reported by the
    GPU in the same run of steps as by the oracle."""
    P = _pyawcm()
    rng = np.random.default_rng(100 + seed)
    L, tf = 7, 0.2
    q, dx = sod_setup(L)
    pts = O.print_times(tf, 41)
    g = P.Solver(1, L, 17, 1.4, dx)
    g.set_schedule(pts)
    o = O.Solver1D("orc", L, 17, 1.4, dx)
    g.upload_q(q)
    o.set_q(q)
    ops = []
    for _ in range(16):
        k = int(rng.integers(0, 5))
        if k == 0:
            ops.append("prepare")
            g.prepare(1e-4)
            o.prepare(1e-4)
        elif k in (1, 2):
            n = int(rng.integers(1, 9))
            ops.append(f"run {n}")
            try:
                steps, _ = g.run(1e-4, 0.5, tf, n)
            except P.AwcmError as e:
                assert "ENONFINITE" in str(e), ops
                broke = False
                for _ in range(n):
                    dt = o.step(1e-4, 0.5, pts)
                    if not (np.isfinite(dt) and np.isfinite(o.get_q()).all()):
                        broke = True
                        break
                assert broke, f"seed {seed}: GPU broke down after {ops}, the oracle did not"
                return
            for _ in range(steps):
                o.step(1e-4, 0.5, pts)
            assert g.t == o.t[0] and g.cnt == o.cnt.value, ops
        elif k == 3:
            ops.append("upload perturbed")
            qn = g.download_q() * (1.0 + 1e-3 * rng.standard_normal(g.npts * g.ncomp))
            g.upload_q(qn)
            o.set_q(qn)
        else:
            ops.append("mask0 pushed back")
            m0 = o.masks()[0][1].copy()
            g.set_mask0(m0)
            o.set_mask0(m0)
        mg, lg = g.masks()
        mo, lo = o.masks()
        assert_same(g.download_q(), o.get_q(), mg, mo, lg, lo, f"seed {seed} after {ops}")


def test_batch_matches_oracle():
    P = _pyawcm()
    rng = np.random.default_rng(7)
    params, qs = [], []
    for k in range(6):
        L = 6 + k % 3
        nx = 16 * 2 ** (L - 1) + 1
        rl, rr = rng.uniform(0.1, 1.0, 2)
        pl, pr = rng.uniform(0.05, 1.0, 2)
        u = rng.uniform(-0.5, 0.5)
        x0 = rng.uniform(0.3, 0.7)
        q, dx, _ = O.ic1(0, nx, 0.0, 1.0, x0, (rl, u, pl), (rr, u, pr))
        params.append((L, 17, 1.4, dx, 1e-3 if k % 2 else 1e-4, 0.5, 0.02))
        qs.append((q, dx))
    pts = O.print_times(0.02, 5)
    b = P.Batch1D(params, pts)
    for k, (q, _) in enumerate(qs):
        b.upload_q(k, q)
    b.run()
    for k, (q, dx) in enumerate(qs):
        L, nxc, g, dx_, tol, cfl, tf = params[k]
        o = O.Solver1D("orc", L, nxc, g, dx_)
        o.set_q(q)
        n = 0
        while o.t[0] < tf:
            o.step(tol, cfl, pts)
            n += 1
        st = b.state(k)
        assert st["status"] == 0 and st["steps"] == n and st["t"] == o.t[0]
        mg, lg = b.masks(k)
        mo, lo = o.masks()
        assert_same(b.download_q(k), o.get_q(), mg, mo, lg, lo, f"batch instance {k}")
