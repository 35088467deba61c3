"""Pins for the CPU oracle (runs without a GPU).

* golden vectors made by the reference build (tests/golden/make_golden.py):
  the oracle must reproduce them bit for bit;
* the reference build itself (oracle/_ref), where present: per-function and
  trajectory comparisons, initial conditions;
* closed forms and invariants that a dropped term, a wrong sign or a
  transposed index would break: dyadic wavelet weights, WENO-Z ideal-weight
  limit and mirror symmetry, flux consistency F(Q,...,Q) = F(Q) (which needs
  R L = I), nested masks, the always-active band, eps = 0 reducing the AWCM
  step to a uniform-grid WENO-Z SSP-RK3 step, Sod against the exact
  Riemann solution, InterpolateAllGridEuler reproducing bicubic fields
  and leaving mask_from's points alone, and the scalar path (AWCM_2D::Prepare,
  WAVELET_2D::Interpolate) agreeing with the Euler path on component 0.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(not O.have_ref(), reason="reference build (oracle/_ref) not present")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ---------------------------------------------------------------- golden ----
def test_oracle_matches_golden_kats():
    g = load("kat_functions.npz")
    L = O.load("orc")
    w = np.zeros(20)
    L.orc_weights(O.dptr(w))
    assert np.array_equal(bits(w.reshape(4, 5)), bits(g["weights"]))
    left = np.array([L.orc_wenoz_left(*r) for r in g["wenoz_in"]])
    right = np.array([L.orc_wenoz_right(*r) for r in g["wenoz_in"]])
    assert np.array_equal(bits(left), bits(g["wenoz_left"]))
    assert np.array_equal(bits(right), bits(g["wenoz_right"]))
    for k in range(len(g["flux1_alpha"])):
        st = np.ascontiguousarray(g["flux1_states"][k]).ravel()
        f = np.zeros(3)
        L.orc_flux1(1.4, O.dptr(st), g["flux1_alpha"][k], O.dptr(f))
        assert np.array_equal(bits(f), bits(g["flux1_out"][k]))
        L.orc_flux1_js(1.4, O.dptr(st), g["flux1_alpha"][k], O.dptr(f))
        assert np.array_equal(bits(f), bits(g["flux1_js_out"][k])), "WENO5-JS flux"
    for k in range(len(g["flux2_alpha"])):
        st = np.ascontiguousarray(g["flux2_states"][k]).ravel()
        for d in range(2):
            f = np.zeros(4)
            L.orc_flux2(1.4, d, O.dptr(st), g["flux2_alpha"][k], O.dptr(f))
            assert np.array_equal(bits(f), bits(g["flux2_out"][k, d]))


@pytest.mark.parametrize("name", ["sod_1d_L6.npz", "shu_osher_1d_L6.npz"])
def test_oracle_matches_golden_1d(name):
    g = load(name)
    s = O.Solver1D("orc", int(g["levels"]), int(g["nx_coarse"]), 1.4, float(g["dx"]))
    s.set_q(g["q0"])
    n = 0
    for target in g["snaps"]:
        while n < target:
            s.step(float(g["tol"]), float(g["cfl"]), g["print_times"])
            n += 1
        m, lv = s.masks()
        assert np.array_equal(m, g[f"masks_{target}"]), f"{name} masks at step {target}"
        assert np.array_equal(lv, g[f"levels_{target}"]), f"{name} levels at step {target}"
        assert np.array_equal(bits(s.get_q()), bits(g[f"q_{target}"])), f"{name} Q at step {target}"
        assert s.t[0] == g[f"t_{target}"] and s.cnt.value == g[f"cnt_{target}"]


@pytest.mark.parametrize("name", ["quad_2d_L4.npz", "dmr_2d_L4.npz"])
def test_oracle_matches_golden_2d(name):
    g = load(name)
    s = O.Solver2D("orc", int(g["levels"]), int(g["nx_coarse"]), int(g["ny_coarse"]), 1.4, int(g["bc"]),
                   0.0, float(g["xu"]), 0.0, 1.0)
    s.set_q(g["q0"])
    n = 0
    for target in g["snaps"]:
        while n < target:
            s.step(float(g["tol"]), float(g["cfl"]), g["print_times"])
            n += 1
        m, lv = s.masks()
        assert np.array_equal(m, g[f"masks_{target}"].reshape(m.shape)), f"{name} masks at step {target}"
        assert np.array_equal(lv, g[f"levels_{target}"].reshape(lv.shape)), f"{name} levels at step {target}"
        assert np.array_equal(bits(s.get_q()), bits(g[f"q_{target}"])), f"{name} Q at step {target}"
        assert s.t[0] == g[f"t_{target}"]


@pytest.mark.parametrize("name", ["all_grid_quad_2d_L4.npz", "all_grid_dmr_2d_L3.npz"])
def test_oracle_all_grid_matches_golden(name):
    """WAVELET_2D::InterpolateAllGridEuler from four masks on a stepped state,
    and two steps after the driver's call (a field that is non-zero outside
    mask0 going into the next PrepareEuler)."""
    g = load(name)
    args = (int(g["levels"]), int(g["nx_coarse"]), int(g["ny_coarse"]), 1.4, int(g["bc"]), 0.0, float(g["xu"]), 0.0, 1.0)
    s = O.Solver2D("orc", *args)
    for f in ("mask0", "mask2", "random", "none"):
        s.set_q(g["q_in"])
        s.interpolate_all(g[f"from_{f}"])
        assert np.array_equal(bits(s.get_q()), bits(g[f"q_{f}"])), f"{name} from {f}"
    s = O.Solver2D("orc", *args)
    s.set_q(g["q0"])
    for _ in range(int(g["steps"])):
        s.step(1e-3, 0.4, g["print_times"])
    assert np.array_equal(bits(s.get_q()), bits(g["q_in"])) and s.t[0] == g["t_in"]
    s.interpolate_all(s.masks()[0][1].copy())
    for _ in range(2):
        s.step(1e-3, 0.4, g["print_times"])
    m, lv = s.masks()
    assert np.array_equal(m, g["masks_after"].reshape(m.shape)) and np.array_equal(lv, g["levels_after"].reshape(lv.shape))
    assert np.array_equal(bits(s.get_q()), bits(g["q_after"])) and s.t[0] == g["t_after"]


def test_oracle_scalar_path_matches_golden():
    """AWCM_2D::Prepare on a scalar field (three rounds) and
    WAVELET_2D::Interpolate from/to four mask pairs, against the reference
    build's vectors."""
    g = load("scalar_box_2d_L4.npz")
    L, nxc = int(g["levels"]), int(g["nx_coarse"])
    n = (nxc - 1) * 2 ** (L - 1) + 1
    s = O.Solver2D("orc", L, nxc, nxc, 1.4, 0, 0.0, 1.0, 0.0, 1.0)
    cur = g["q0"]
    for r in range(3):
        s.set_scalar(cur)
        s.prepare_scalar(1e-4)
        m, lv = s.masks()
        assert np.array_equal(m, g[f"masks_{r}"].reshape(m.shape)), f"masks, round {r}"
        assert np.array_equal(lv, g[f"levels_{r}"].reshape(lv.shape)), f"levels, round {r}"
        assert np.array_equal(bits(s.get_scalar()), bits(g[f"q_{r}"])), f"field, round {r}"
        cur = (np.roll(s.get_scalar().reshape(n + 6, n + 6), 1, axis=1) * 0.999).ravel()
    for name in ("m0_all", "m0_m3", "m1_m2", "random"):
        s.set_scalar(g["q_2"])
        s.interpolate_scalar(g[f"from_{name}"], g[f"to_{name}"])
        assert np.array_equal(bits(s.get_scalar()), bits(g[f"qi_{name}"])), name


def test_oracle_scalar_1d_matches_golden():
    """AWCM_1D::Prepare on a scalar field (three rounds) and
    WAVELET_1D::Interpolate (both forms) against the reference build's
    vectors."""
    g = load("scalar_1d_L7.npz")
    L, nxc = int(g["levels"]), int(g["nx_coarse"])
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    s = O.Solver1D("orc", L, nxc, 1.4, 1.0 / nx)
    cur = g["q0"]
    for r in range(3):
        s.set_scalar(cur)
        s.prepare_scalar(1e-4)
        m, lv = s.masks()
        assert np.array_equal(m, g[f"masks_{r}"].reshape(m.shape)), f"masks, round {r}"
        assert np.array_equal(lv, g[f"levels_{r}"].reshape(lv.shape)), f"levels, round {r}"
        assert np.array_equal(bits(s.get_scalar()), bits(g[f"q_{r}"])), f"field, round {r}"
        cur = np.roll(s.get_scalar(), 2) * 0.999
    for name in ("m0_all", "m0_m3", "m1_m2", "random"):
        s.set_scalar(g["q_2"])
        s.interpolate_scalar(g[f"from_{name}"], g[f"to_{name}"] if f"to_{name}" in g.files else None)
        assert np.array_equal(bits(s.get_scalar()), bits(g[f"qi_{name}"])), name


def test_oracle_euler_interp_1d_matches_golden():
    """WAVELET_1D::InterpolateEuler, both forms, against the reference build's
    vectors."""
    g = load("euler_interp_1d_L6.npz")
    s = O.Solver1D("orc", int(g["levels"]), int(g["nx_coarse"]), 1.4, float(g["dx"]))
    for name in ("m0_all", "m0_m3", "m1_m2", "random_all"):
        s.set_q(g["q_in"])
        s.interpolate_euler(g[f"from_{name}"], g[f"to_{name}"] if f"to_{name}" in g.files else None)
        assert np.array_equal(bits(s.get_q()), bits(g[f"q_{name}"])), name


def test_scalar_1d_interpolation_reproduces_cubics():
    """WAVELET_1D::Interpolate(q, mra, from) with from = the coarsest grid and
    both ends (the points no lattice reaches): a cubic comes back everywhere
    to rounding from garbage."""
    L, nxc = 6, 9
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    x = np.arange(-3, nx + 3) / (nx - 1)
    exact = 0.3 - 1.2 * x + 2.5 * x ** 2 - 1.7 * x ** 3
    f = np.zeros(nx + 6, np.uint8)
    f[3:-3:2 ** (L - 1)] = 1
    f[:4] = 1
    f[-4:] = 1
    rng = np.random.default_rng(3)
    q = np.where(f == 1, exact, rng.uniform(-20, 20, exact.shape))
    s = O.Solver1D("orc", L, nxc, 1.4, 1.0 / nx)
    s.set_scalar(q)
    s.interpolate_scalar(f)
    assert np.abs(s.get_scalar()[3:-3] - exact[3:-3]).max() < 1e-12


# ----------------------------------------------------- reference build ----
@needs_ref
def test_oracle_ics_match_reference():
    for kind, res in ((0, 257), (1, 513)):
        q_ref = np.zeros((res + 6) * 3)
        dx_r, x0_r = np.zeros(1), np.zeros(1)
        O.load("ref").ref_ic1(kind, res, O.dptr(q_ref), O.dptr(dx_r), O.dptr(x0_r))
        if kind == 0:
            q, dx, x0 = O.ic1(0, res, 0.0, 10.0, 5.0, O.SOD_L, O.SOD_R)
        else:
            q, dx, x0 = O.ic1(1, res, -5.0, 5.0, -4.0, O.SHU_L, O.SHU_L)
        assert np.array_equal(bits(q), bits(q_ref)) and dx == dx_r[0] and x0 == x0_r[0]
    for kind, res in ((0, 65), (1, 33)):
        nx = res if kind == 0 else (res - 1) * 4 + 1
        q_ref = np.zeros((nx + 6) * (res + 6) * 4)
        dx_r, dy_r = np.zeros(1), np.zeros(1)
        O.load("ref").ref_ic2(kind, res, O.dptr(q_ref), O.dptr(dx_r), O.dptr(dy_r))
        if kind == 0:
            q, dx, _ = O.ic2(0, nx, res, 0.0, 1.0, 0.0, 1.0, 0.8)
        else:
            q, dx, _ = O.ic2(1, nx, res, 0.0, 4.0, 0.0, 1.0, 1.5)  # the reference IC uses rho_pre = 1.5
        assert np.array_equal(bits(q), bits(q_ref)) and dx == dx_r[0]


@needs_ref
def test_oracle_functions_match_reference_random():
    rng = np.random.default_rng(11)
    A, B = O.load("orc"), O.load("ref")
    for _ in range(200):
        r = rng.normal(size=5) * rng.choice([1e-6, 1.0, 1e3])
        assert bits(A.orc_wenoz_left(*r)) == bits(B.ref_wenoz_left(*r))
        assert bits(A.orc_wenoz_right(*r)) == bits(B.ref_wenoz_right(*r))


@needs_ref
def test_oracle_sod_full_run_matches_reference():
    L, nxc = 7, 17
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    a, b = O.Solver1D("orc", L, nxc, 1.4, dx), O.Solver1D("ref", L, nxc, 1.4, dx)
    a.set_q(q)
    b.set_q(q)
    pts = O.print_times(0.2, 41)
    while a.t[0] < 0.2:
        assert a.step(1e-4, 0.5, pts) == b.step(1e-4, 0.5, pts)
    ma, la = a.masks()
    mb, lb = b.masks()
    assert np.array_equal(ma, mb) and np.array_equal(la, lb)
    assert np.array_equal(bits(a.get_q()), bits(b.get_q()))
    assert O.load("orc").orc1_oob(a.h) == 0


@needs_ref
def test_oracle_sod_weno5_full_run_matches_reference():
    """The shipped 1-D driver's stage flux, WENO_EULER_SOLVER_1D::ComputedF
    (main_1d_euler.cpp:109): per-face known answers on random states, and a
    full Sod run (918 steps) bit for bit."""
    A, B = O.load("orc"), O.load("ref")
    rng = np.random.default_rng(17)
    for _ in range(200):
        rho, u, p = rng.uniform(0.1, 2, 6), rng.uniform(-1.5, 1.5, 6), rng.uniform(0.05, 2, 6)
        st = np.stack([rho, rho * u, p / 0.4 + 0.5 * rho * u * u], 1).ravel()
        a = rng.uniform(0.5, 3)
        fa, fb = np.zeros(3), np.zeros(3)
        A.orc_flux1_js(1.4, O.dptr(st), a, O.dptr(fa))
        B.ref_flux1_js(1.4, O.dptr(st), a, O.dptr(fb))
        assert np.array_equal(bits(fa), bits(fb))
    L, nxc = 7, 17
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    a, b = O.Solver1D("orc", L, nxc, 1.4, dx), O.Solver1D("ref", L, nxc, 1.4, dx)
    a.set_flux(1)
    b.set_flux(1)
    a.set_q(q)
    b.set_q(q)
    pts = O.print_times(0.2, 41)
    while a.t[0] < 0.2:
        assert a.step(1e-4, 0.5, pts) == b.step(1e-4, 0.5, pts)
    ma, la = a.masks()
    mb, lb = b.masks()
    assert np.array_equal(ma, mb) and np.array_equal(la, lb)
    assert np.array_equal(bits(a.get_q()), bits(b.get_q()))


@needs_ref
@pytest.mark.parametrize("case", ["quad", "dmr"])
def test_oracle_2d_matches_reference(case):
    if case == "quad":
        L, nxc, nyc, bc, xu, s = 5, 9, 9, 0, 1.0, 0.5
    else:
        L, nxc, nyc, bc, xu, s = 4, 33, 9, 1, 4.0, 1.4
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    ny = (nyc - 1) * 2 ** (L - 1) + 1
    q, dx, _ = O.ic2(bc, nx, ny, 0.0, xu, 0.0, 1.0, s)
    a = O.Solver2D("orc", L, nxc, nyc, 1.4, bc, 0.0, xu, 0.0, 1.0)
    b = O.Solver2D("ref", L, nxc, nyc, 1.4, bc, 0.0, xu, 0.0, 1.0)
    a.set_q(q)
    b.set_q(q)
    pts = O.print_times(0.25, 9)
    for _ in range(12):
        assert a.step(1e-3, 0.4, pts) == b.step(1e-3, 0.4, pts)
        ma, la = a.masks()
        mb, lb = b.masks()
        assert np.array_equal(ma, mb) and np.array_equal(la, lb)
        assert np.array_equal(bits(a.get_q()), bits(b.get_q()))
    assert O.load("orc").orc2_oob(a.h) == 0


@needs_ref
@pytest.mark.parametrize("case", ["quad", "dmr"])
def test_oracle_all_grid_matches_reference(case):
    if case == "quad":
        L, nxc, nyc, bc, xu, s = 5, 9, 9, 0, 1.0, 0.5
    else:
        L, nxc, nyc, bc, xu, s = 4, 33, 9, 1, 4.0, 1.4
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    ny = (nyc - 1) * 2 ** (L - 1) + 1
    q, dx, _ = O.ic2(bc, nx, ny, 0.0, xu, 0.0, 1.0, s)
    a = O.Solver2D("orc", L, nxc, nyc, 1.4, bc, 0.0, xu, 0.0, 1.0)
    b = O.Solver2D("ref", L, nxc, nyc, 1.4, bc, 0.0, xu, 0.0, 1.0)
    a.set_q(q)
    b.set_q(q)
    pts = O.print_times(0.25, 9)
    rng = np.random.default_rng(7)
    for _ in range(5):
        a.step(1e-3, 0.4, pts)
        b.step(1e-3, 0.4, pts)
    m = a.masks()[0]
    q5 = a.get_q()
    # mask1, mask3 and the random masks hold points whose field the step's
    # residue removal zeroed: their results are not physical states, but
    # they are still the same arithmetic
    for f in (m[2], m[4], (rng.random(m[1].shape) < 0.1).astype(np.uint8),
              (rng.random(m[1].shape) < 0.7).astype(np.uint8), m[1]):
        a.set_q(q5)
        b.set_q(q5)
        a.interpolate_all(f)
        b.interpolate_all(f)
        assert np.array_equal(bits(a.get_q()), bits(b.get_q()))
    # the driver's call (from mask0), then the run goes on from that field
    for _ in range(3):
        assert a.step(1e-3, 0.4, pts) == b.step(1e-3, 0.4, pts)
        ma, la = a.masks()
        mb, lb = b.masks()
        assert np.array_equal(ma, mb) and np.array_equal(la, lb)
        assert np.array_equal(bits(a.get_q()), bits(b.get_q()))
    assert O.load("orc").orc2_oob(a.h) == 0


@needs_ref
@pytest.mark.parametrize("case", ["box", "dmr_density"])
def test_oracle_scalar_path_matches_reference(case):
    """AWCM_2D::Prepare (four rounds, the field moved between them),
    WAVELET_2D::Interpolate and InterpolateEuler from/to five mask pairs."""
    if case == "box":
        L, nxc, nyc, bc, xu = 5, 9, 9, 0, 1.0
    else:
        L, nxc, nyc, bc, xu = 4, 33, 9, 1, 4.0
    nx, ny = (nxc - 1) * 2 ** (L - 1) + 1, (nyc - 1) * 2 ** (L - 1) + 1
    a = O.Solver2D("orc", L, nxc, nyc, 1.4, bc, 0.0, xu, 0.0, 1.0)
    b = O.Solver2D("ref", L, nxc, nyc, 1.4, bc, 0.0, xu, 0.0, 1.0)
    if case == "box":
        x = (np.arange(-3, nx + 3) + 0.5) / nx
        X, Y = np.meshgrid(x, x)
        q = (np.exp(-50 * (X - 0.5) ** 2 - 50 * (Y - 0.5) ** 2) + 0.3 * (X > 0.7)).ravel()
    else:
        q = O.ic2(1, nx, ny, 0.0, 4.0, 0.0, 1.0, 1.4)[0].reshape(-1, 4)[:, 0].copy()
    cur = q
    for r in range(4):
        a.set_scalar(cur)
        b.set_scalar(cur)
        a.prepare_scalar(1e-4)
        b.prepare_scalar(1e-4)
        ma, la = a.masks()
        mb, lb = b.masks()
        assert np.array_equal(ma, mb) and np.array_equal(la, lb), f"round {r}"
        assert np.array_equal(bits(a.get_scalar()), bits(b.get_scalar())), f"round {r}"
        cur = (np.roll(a.get_scalar().reshape(ny + 6, nx + 6), 1, axis=1) * 0.999).ravel()
    rng = np.random.default_rng(13)
    m = ma
    qe = np.stack([q, 2 * q, q - 1, 3 * q + 1], axis=-1).ravel()
    for f, t in ((m[1], np.ones_like(m[1])), (m[1], m[4]), (m[2], m[3]), (np.zeros_like(m[1]), m[4]),
                 ((rng.random(m[1].shape) < 0.3).astype(np.uint8), (rng.random(m[1].shape) < 0.6).astype(np.uint8))):
        a.set_scalar(q)
        b.set_scalar(q)
        a.interpolate_scalar(f, t)
        b.interpolate_scalar(f, t)
        assert np.array_equal(bits(a.get_scalar()), bits(b.get_scalar()))
        a.set_q(qe)
        b.set_q(qe)
        a.interpolate_euler(f, t)
        b.interpolate_euler(f, t)
        assert np.array_equal(bits(a.get_q()), bits(b.get_q()))


@pytest.mark.parametrize("name", ["small", "driver"])
def test_oracle_advect_matches_reference_driver_text(name):
    """The box-advection driver (Wavelet/main_2d_advect.cpp:46-199) on the
    oracle against the reference build's own run of it
    (tests/golden/api_probe_advect.txt): after every Prepare and every
    iteration, the hashes of the ghosted field, the five masks and level0..2,
    over 40 iterations of 65 x 65 and 10 of the driver's 449 x 449."""
    import advect_lib as A
    gold = A.golden()
    L, nxc, eps, steps, res, q, dx, dt = A.setup(name)
    xl, xu, _ = O.box_grid(res)
    S = O.Solver2D("orc", L, nxc, nxc, 1.4, 0, xl, xu, xl, xu)
    S.set_scalar(q)
    for s in range(steps):
        S.prepare_scalar(eps)
        m, lv = S.masks()
        assert A.report(S.get_scalar(), m, lv) == gold[f"{name}.{s}.prepare"], s
        S.advect(eps, dt, 1, with_prepare=False)
        m, lv = S.masks()
        assert A.report(S.get_scalar(), m, lv) == gold[f"{name}.{s}.step"], s
    assert O.load("orc").orc2_oob(S.h) == 0


@needs_ref
def test_oracle_scalar_1d_matches_reference():
    """AWCM_1D::Prepare (four rounds) and WAVELET_1D::Interpolate (both forms)
    against the reference build, on two grids."""
    for L, nxc in ((7, 17), (9, 16)):
        nx = (nxc - 1) * 2 ** (L - 1) + 1
        x = (np.arange(-3, nx + 3) + 0.5) / nx
        q = np.exp(-200 * (x - 0.4) ** 2) + 0.5 * (np.abs(x - 0.7) < 0.1)
        a, b = O.Solver1D("orc", L, nxc, 1.4, 1.0 / nx), O.Solver1D("ref", L, nxc, 1.4, 1.0 / nx)
        cur = q
        for r in range(4):
            a.set_scalar(cur)
            b.set_scalar(cur)
            a.prepare_scalar(1e-4)
            b.prepare_scalar(1e-4)
            ma, la = a.masks()
            mb, lb = b.masks()
            assert np.array_equal(ma, mb) and np.array_equal(la, lb), (L, r)
            assert np.array_equal(bits(a.get_scalar()), bits(b.get_scalar())), (L, r)
            cur = np.roll(a.get_scalar(), 2) * 0.999
        rng = np.random.default_rng(2)
        m = ma
        for f, t in ((m[1], None), (m[1], m[4]), (m[2], m[3]),
                     ((rng.random(m[1].shape) < 0.3).astype(np.uint8), (rng.random(m[1].shape) < 0.6).astype(np.uint8)),
                     ((rng.random(m[1].shape) < 0.2).astype(np.uint8), None)):
            a.set_scalar(q)
            b.set_scalar(q)
            a.interpolate_scalar(f, t)
            b.interpolate_scalar(f, t)
            assert np.array_equal(bits(a.get_scalar()), bits(b.get_scalar()))


# ----------------------------------------------- closed forms, invariants ----
def test_wavelet_weights_closed_form():
    w = np.zeros(20)
    O.load("orc").orc_weights(O.dptr(w))
    w = w.reshape(4, 5)
    # weight(i, j) at w[i, j+1]: Lagrange basis of nodes 0..3 at j + 1/2
    assert np.array_equal(w[:, 1] * 16, [5, 15, -5, 1])    # j = 0: left-boundary predictor
    assert np.array_equal(w[:, 2] * 16, [-1, 9, 9, -1])    # j = 1: centred predictor
    assert np.array_equal(w[:, 3] * 16, [1, -5, 15, 5])    # j = 2: right-boundary predictor
    for j in range(5):
        assert abs(w[:, j].sum() - 1.0) < 1e-15          # partition of unity
        x = j - 1 + 0.5
        assert abs((w[:, j] * np.arange(4) ** 3).sum() - x ** 3) < 1e-12  # exact for cubics


def test_wenoz_ideal_limit_and_mirror():
    L = O.load("orc")
    rng = np.random.default_rng(3)
    for _ in range(100):
        a, b = rng.normal(size=2)
        v = a + b * np.arange(-2, 3)          # linear data: all smoothness indicators equal
        left = L.orc_wenoz_left(*v)
        assert abs(left - (a - 0.5 * b)) <= 1e-13 * (1 + abs(a) + abs(b))
        r = rng.normal(size=5)
        assert bits(L.orc_wenoz_right(*r)) == bits(L.orc_wenoz_left(*r[::-1]))
    # smooth data: weights near ideal, value near the 5th-order linear scheme
    x = 0.01 * np.arange(-2, 3)
    v = np.sin(x) + 2
    lin = (-3 * v[0] + 27 * v[1] + 47 * v[2] - 13 * v[3] + 2 * v[4]) / 60
    assert abs(L.orc_wenoz_left(*v) - lin) < 1e-9


def _phys_flux(q, dim, d=0, g=1.4):
    if dim == 1:
        rho, m, E = q
        u = m / rho
        p = (g - 1) * (E - 0.5 * m * m / rho)
        return np.array([m, m * u + p, (E + p) * u])
    rho, mx, my, E = q
    u, v = mx / rho, my / rho
    p = (g - 1) * (E - 0.5 * (mx * mx + my * my) / rho)
    vn = u if d == 0 else v
    return np.array([rho * vn, mx * vn + (p if d == 0 else 0), my * vn + (p if d == 1 else 0), (E + p) * vn])


def test_face_flux_consistency():
    L = O.load("orc")
    rng = np.random.default_rng(5)
    for _ in range(50):
        rho, u, v, p = rng.uniform(0.2, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.2, 2)
        q1 = np.array([rho, rho * u, p / 0.4 + 0.5 * rho * u * u])
        st = np.tile(q1, 6)
        f = np.zeros(3)
        a = rng.uniform(1, 3)
        L.orc_flux1(1.4, O.dptr(st), a, O.dptr(f))
        assert np.allclose(f, _phys_flux(q1, 1), rtol=1e-12, atol=1e-12)
        L.orc_flux1_js(1.4, O.dptr(st), a, O.dptr(f))
        assert np.allclose(f, _phys_flux(q1, 1), rtol=1e-12, atol=1e-12), "WENO5-JS flux"
        q2 = np.array([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])
        st = np.tile(q2, 6)
        for d in range(2):
            f = np.zeros(4)
            L.orc_flux2(1.4, d, O.dptr(st), rng.uniform(1, 3), O.dptr(f))
            assert np.allclose(f, _phys_flux(q2, 2, d), rtol=1e-12, atol=1e-12)


def test_mask_invariants_2d():
    L, nxc, nyc = 4, 33, 9
    nx, ny = (nxc - 1) * 8 + 1, (nyc - 1) * 8 + 1
    q, dx, _ = O.ic2(1, nx, ny, 0.0, 4.0, 0.0, 1.0, 1.4)
    s = O.Solver2D("orc", L, nxc, nyc, 1.4, 1, 0.0, 4.0, 0.0, 1.0)
    s.set_q(q)
    pts = O.print_times(0.2, 21)
    for _ in range(4):
        s.step(1e-3, 0.4, pts)
        m, lv = s.masks()
        inner = m[:, 3:-3, 3:-3].astype(bool)
        for k in range(1, 4):
            assert not (inner[k] & ~inner[k + 1]).any(), "mask nesting"
        for k in range(1, 5):
            band = np.ones_like(inner[k])
            band[2:-2, 2:-2] = False
            assert inner[k][band].all(), "2-wide boundary band"
            m0 = 2 ** (L - 1)
            assert inner[k][::m0, ::m0].all(), "coarsest grid always active"
        # level == max wherever a +-1 neighbour is active (level0 from mask0)
        act = np.pad(inner[1], 1)
        nb = act[:-2, 1:-1] | act[2:, 1:-1] | act[1:-1, :-2] | act[1:-1, 2:]
        assert (lv[0][inner[1] & nb] == L - 1).all()
        assert (lv[0][~inner[1]] == 0).all()


def _bicubic_fields(nx, ny, rng):
    """Four fields, each a random polynomial of degree <= 3 in x and in y."""
    x = np.arange(-3, nx + 3) / (nx - 1)
    y = np.arange(-3, ny + 3) / (ny - 1)
    X, Y = np.meshgrid(x, y)
    out = np.zeros((ny + 6, nx + 6, 4))
    for p in range(4):
        c = rng.uniform(-1.0, 1.0, (4, 4))
        out[:, :, p] = sum(c[a, b] * X ** a * Y ** b for a in range(4) for b in range(4))
    return out


@pytest.mark.parametrize("bc,L,nxc,nyc", [(0, 5, 5, 5), (1, 4, 9, 5)])
def test_all_grid_reproduces_bicubic_fields(bc, L, nxc, nyc):
    """The 4-point interpolating wavelet (and its one-sided forms at the
    walls) is exact for cubics, and its d3 prediction P_x + P_y - P_xy is exact
    for products of cubics: a bicubic field known only on the coarsest grid
    and on the outer ring of the finest grid (the boundary layer no sub-band
    lattice reaches, MRA_GRID_UNIFORM_2D.h:102-124) comes back everywhere to
    rounding, in all four components.  The other points start from garbage,
    so a component, sub-band or level that is skipped, a transposed operand
    or a wrong stencil index fails."""
    rng = np.random.default_rng(5 + bc)
    nx, ny = (nxc - 1) * 2 ** (L - 1) + 1, (nyc - 1) * 2 ** (L - 1) + 1
    s = O.Solver2D("orc", L, nxc, nyc, 1.4, bc, 0.0, 1.0, 0.0, 1.0)
    exact = _bicubic_fields(nx, ny, rng)
    m = 2 ** (L - 1)
    coarse = np.zeros((ny + 6, nx + 6), dtype=np.uint8)
    coarse[3:-3:m, 3:-3:m] = 1
    work = np.zeros(coarse.shape, bool)  # points a sub-band lattice reaches
    work[4:-4, 4:-4] = True
    for f in (coarse, coarse | (rng.random(coarse.shape) < 0.2)):
        q = np.where((work & (f == 0))[:, :, None], rng.uniform(-50, 50, exact.shape), exact)
        s.set_q(q.ravel())
        s.interpolate_all(f)
        got = s.get_q().reshape(exact.shape)
        inner = (slice(3, -3), slice(3, -3))
        err = np.abs(got[inner] - exact[inner]).max()
        assert err < 1e-12, f"max error {err}"
        # the outer ring and the ghost cells are never written
        assert np.array_equal(bits(got[~work]), bits(q[~work]))


def test_all_grid_leaves_mask_from_points():
    """Points in mask_from, the coarsest grid (never a detail point) and the
    outer ring of the finest grid keep their bits, every other point gets a
    new value; mask_from covering everything changes nothing."""
    L, nxc, nyc = 4, 17, 6
    nx, ny = (nxc - 1) * 8 + 1, (nyc - 1) * 8 + 1
    rng = np.random.default_rng(9)
    s = O.Solver2D("orc", L, nxc, nyc, 1.4, 1, 0.0, 4.0, 0.0, 1.0)
    q = rng.uniform(0.5, 2.0, (ny + 6) * (nx + 6) * 4)
    for f in (np.ones((ny + 6, nx + 6), np.uint8), (rng.random((ny + 6, nx + 6)) < 0.5).astype(np.uint8),
              np.zeros((ny + 6, nx + 6), np.uint8)):
        s.set_q(q)
        s.interpolate_all(f)
        got = s.get_q().reshape(ny + 6, nx + 6, 4)
        keep = f.astype(bool).copy()
        keep[3:-3:8, 3:-3:8] = True
        assert np.array_equal(bits(got[keep]), bits(q.reshape(got.shape)[keep]))
        changed = (bits(got) != bits(q.reshape(got.shape))).any(axis=2)
        work = np.zeros_like(keep)  # all but the outer ring of the finest grid
        work[4:-4, 4:-4] = True
        assert not changed[~work].any()
        assert changed[work & ~keep].all()


def test_scalar_path_is_the_euler_path_on_component_0():
    """The reference's scalar transforms are its Euler transforms on
    component 0 (WAVELET_2D.h:37-359 against :478-783): Prepare and
    PrepareEuler give the same masks, levels and first component, and
    Interpolate(from, all points) is InterpolateAllGridEuler's first
    component; Interpolate(from, to) changes only points of
    CheckReconstruction(to minus from) outside from."""
    L, nxc, nyc = 4, 33, 9
    nx, ny = (nxc - 1) * 8 + 1, (nyc - 1) * 8 + 1
    q4 = O.ic2(1, nx, ny, 0.0, 4.0, 0.0, 1.0, 1.4)[0]
    a = O.Solver2D("orc", L, nxc, nyc, 1.4, 1, 0.0, 4.0, 0.0, 1.0)
    b = O.Solver2D("orc", L, nxc, nyc, 1.4, 1, 0.0, 4.0, 0.0, 1.0)
    a.set_q(q4)
    a.prepare(1e-3)
    b.set_scalar(q4.reshape(-1, 4)[:, 0])
    b.prepare_scalar(1e-3)
    ma, la = a.masks()
    mb, lb = b.masks()
    assert np.array_equal(ma[1:], mb[1:]) and np.array_equal(la[:3], lb[:3])
    assert not mb[0].any(), "the scalar Prepare leaves mask_temp alone"
    assert np.array_equal(bits(a.get_q().reshape(-1, 4)[:, 0]), bits(b.get_scalar()))
    q = a.get_q()
    f = ma[1]
    a.set_q(q)
    a.interpolate_all(f)
    b.set_scalar(q.reshape(-1, 4)[:, 0])
    b.interpolate_scalar(f, np.ones_like(f))
    assert np.array_equal(bits(a.get_q().reshape(-1, 4)[:, 0]), bits(b.get_scalar()))
    # a `to` of one detail point far from the walls: only that point moves
    to = np.zeros_like(f)
    j, i = 3 + 8 * 3 + 4, 3 + 8 * 10 + 4  # a d3 point of the level of spacing 8
    to[j, i] = 1
    f2 = np.zeros_like(f)
    b.set_scalar(q.reshape(-1, 4)[:, 0])
    b.interpolate_scalar(f2, to)
    moved = b.get_scalar().reshape(ny + 6, nx + 6) != q.reshape(ny + 6, nx + 6, 4)[:, :, 0]
    assert moved[j, i] and moved.sum() >= 1
    # the closure of one d3 point only adds the coarser points its stencils need
    assert not moved[:, :3].any() and not moved[:, -3:].any()


def _uniform_step_1d(q, dx, cfl, g=1.4):
    """Textbook uniform-grid WENO-Z5/Roe SSP-RK3 step with the drivers' operation order."""
    L = O.load("orc")
    n = q.size // 3 - 6
    Q = q.reshape(-1, 3).copy()

    def bc(A):
        A[0:3] = A[3]
        A[-3:] = A[-4]

    def sr(A):
        rho, m, E = A[3:-3, 0], A[3:-3, 1], A[3:-3, 2]
        u = m / rho
        p = (g - 1) * (E - 0.5 * (m * m) / rho)
        return np.abs(u) + np.sqrt(g * p / rho)

    def dF(A, alpha):
        F = np.zeros((n + 1, 3))
        for f in range(n + 1):
            st = np.ascontiguousarray(A[f:f + 6]).ravel()  # points f-3 .. f+2
            L.orc_flux1(g, O.dptr(st), alpha, O.dptr(F[f]))
        return F[1:] - F[:-1]

    vmax = 0.0
    for v in sr(Q):
        vmax = v if v > vmax else vmax
    dt = cfl * dx / vmax
    h = dt / (dx * 1)
    bc(Q)
    Qc = Q.copy()
    Q[3:-3] = Q[3:-3] - h * dF(Q, sr(Q).max())
    bc(Q)
    Q[3:-3] = 0.75 * Qc[3:-3] + 0.25 * (Q[3:-3] - h * dF(Q, sr(Q).max()))
    bc(Q)
    Q[3:-3] = (Qc[3:-3] + 2.0 * (Q[3:-3] - h * dF(Q, sr(Q).max()))) / 3.0
    return Q.ravel(), dt


def test_eps0_reduces_to_uniform_wenoz():
    """With eps = 0 nothing is thresholded: every point is active at the finest
    level and the AWCM step is the uniform WENO-Z SSP-RK3 step (up to the
    rounding of the forward/inverse transform round trip)."""
    L, nxc = 5, 9
    nx = (nxc - 1) * 16 + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    s = O.Solver1D("orc", L, nxc, 1.4, dx)
    s.set_q(q)
    dt = s.step(0.0, 0.5, np.array([0.0, 1.0]))
    m, lv = s.masks()
    assert m[1:, 3:-3].all() and (lv == L - 1).all()
    qu, dtu = _uniform_step_1d(q, dx, 0.5)
    assert abs(dt - dtu) <= 1e-15 * dtu
    a, b = s.get_q().reshape(-1, 3)[3:-3], qu.reshape(-1, 3)[3:-3]
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def _exact_sod(x, t, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1), x0=0.5, g=1.4):
    """Exact Riemann solution (Toro, ch. 4) sampled at x, t."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)

    def f(p, rk, pk, ck):
        if p > pk:
            A, B = 2 / ((g + 1) * rk), (g - 1) / (g + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B))
        return 2 * ck / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)

    lo, hi = 1e-8, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid, rl, pl, cl) + f(mid, rr, pr, cr) + ur - ul > 0:
            hi = mid
        else:
            lo = mid
    ps = 0.5 * (lo + hi)
    us = 0.5 * (ul + ur) + 0.5 * (f(ps, rr, pr, cr) - f(ps, rl, pl, cl))
    rho = np.empty_like(x)
    for k, xi in enumerate(x):
        s = (xi - x0) / t
        if s < us:  # left of contact: rarefaction (Sod)
            shl, stl = ul - cl, us - cl * (ps / pl) ** ((g - 1) / (2 * g))
            if s < shl:
                rho[k] = rl
            elif s > stl:
                rho[k] = rl * (ps / pl) ** (1 / g)
            else:
                rho[k] = rl * (2 / (g + 1) + (g - 1) / ((g + 1) * cl) * (ul - s)) ** (2 / (g - 1))
        else:  # right: shock
            sh = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
            rho[k] = rr if s > sh else rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
    return rho


def test_sod_against_exact_riemann_solution():
    L, nxc = 7, 17
    nx = (nxc - 1) * 64 + 1
    q, dx, x0 = O.ic1(0, nx, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    s = O.Solver1D("orc", L, nxc, 1.4, dx)
    s.set_q(q)
    pts = O.print_times(0.2, 41)
    while s.t[0] < 0.2:
        s.step(1e-4, 0.5, pts)
    m, _ = s.masks()
    act = m[1, 3:-3].astype(bool)
    x = x0 + dx * np.arange(nx)
    rho = s.get_q().reshape(-1, 3)[3:-3, 0]
    ex = _exact_sod(x[act], 0.2)
    assert np.abs(rho[act] - ex).mean() < 5e-3


def test_batched_workload_instance_that_breaks_down():
    """Instance 448 of the seeded batched workload breaks down in the oracle
    (bit-exact with the reference): after 1558 steps, the next PrepareEuler
    leaves a mask3 point with a non-finite wave speed |u| + c.  The
    reference's MAX (a ternary) drops that NaN from the CFL maximum, so its
    time only turns NaN at step 1657; the GPU stops the instance at the first
    non-finite wave speed instead."""
    import workloads
    L, rl, rr, pl, pr, u, x0, eps = workloads.shock_tube_instances(512)[448]
    nx = 16 * 2 ** (L - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, x0, (rl, u, pl), (rr, u, pr))
    o = O.Solver1D("orc", L, 17, 1.4, dx)
    o.set_q(q)
    pts = O.print_times(0.15, 41)

    def nonfinite_speed_after_prepare():
        saved, (m, lv) = o.get_q(), o.masks()
        o.prepare(eps)
        Q = o.get_q().reshape(-1, 3)[3:-3]
        m3 = o.masks()[0][4][3:-3] > 0
        with np.errstate(all="ignore"):
            p = 0.4 * (Q[:, 2] - 0.5 * Q[:, 1] ** 2 / Q[:, 0])
            s = np.abs(Q[:, 1] / Q[:, 0]) + np.sqrt(1.4 * p / Q[:, 0])
        o.set_q(saved)
        o.set_mask0(m[1])
        return not np.all(np.isfinite(s[m3]))

    for n in range(1558):
        o.step(eps, 0.5, pts)
    assert np.isfinite(o.t[0])
    assert nonfinite_speed_after_prepare()
    for n in range(1558, 1657):
        assert np.isfinite(o.t[0]), n
        o.step(eps, 0.5, pts)
    assert not np.isfinite(o.t[0])


@needs_ref
def test_oracle_matches_reference_on_largest_1d_grid():
    """14 levels over 8 coarse points (57345 finest points, the most levels
    and nearly the largest grid the C ABI accepts): 30 steps of a Sod tube,
    oracle and reference build bit-identical in t, masks, levels and state."""
    L, nxc = 14, 8
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    pts = O.print_times(0.2, 41)
    o = O.Solver1D("orc", L, nxc, 1.4, dx)
    r = O.Solver1D("ref", L, nxc, 1.4, dx)
    o.set_q(q)
    r.set_q(q)
    for _ in range(30):
        o.step(1e-4, 0.5, pts)
        r.step(1e-4, 0.5, pts)
    assert o.t[0] == r.t[0] > 0
    mo, lo = o.masks()
    mr, lr = r.masks()
    assert all(np.array_equal(a, b) for a, b in zip(mo, mr)) and np.array_equal(lo, lr)
    assert np.array_equal(o.get_q().view(np.uint64), r.get_q().view(np.uint64))
    # the discontinuities are resolved down to the finest of the 14 levels
    assert lo.max() == L - 1 and len(np.unique(lo)) >= L - 2


@needs_ref
@pytest.mark.parametrize("shape", [(4096, 5), (5, 4096)])
def test_oracle_matches_reference_on_index_limit_strip(shape):
    """A 65521 x 65 strip and its transpose (5 levels, square cells, the
    quadrant problem): rows and columns at the 16-bit limit of the device's
    packed list entries.  Two steps, bit-identical in t, masks, levels and
    state."""
    L, (nxc, nyc) = 5, shape
    nx, ny = (nxc - 1) * 16 + 1, (nyc - 1) * 16 + 1
    h = 1.0 / (max(nx, ny) - 1)
    xu, yu = (nx - 1) * h, (ny - 1) * h
    q, _, _ = O.ic2(0, nx, ny, 0.0, xu, 0.0, yu, 0.5)
    pts = O.print_times(0.25, 9)
    o = O.Solver2D("orc", L, nxc, nyc, 1.4, 0, 0.0, xu, 0.0, yu)
    r = O.Solver2D("ref", L, nxc, nyc, 1.4, 0, 0.0, xu, 0.0, yu)
    o.set_q(q)
    r.set_q(q)
    for _ in range(2):
        o.step(1e-3, 0.4, pts)
        r.step(1e-3, 0.4, pts)
    assert o.t[0] == r.t[0] > 0
    mo, lo = o.masks()
    mr, lr = r.masks()
    assert all(np.array_equal(a, b) for a, b in zip(mo, mr)) and np.array_equal(lo, lr)
    assert np.array_equal(o.get_q().view(np.uint64), r.get_q().view(np.uint64))


def mask_cases_2d(S, rng):
    """Masks from a run (mask0, mask1) and a seeded random one with the 2-wide
    band every AWCM mask carries, which at 4 levels keeps every ComputeLevel
    search inside the ghosts."""
    m, lv = S.masks()
    ny, nx = m.shape[1] - 6, m.shape[2] - 6
    r = (rng.random((ny + 6, nx + 6)) < 0.05).astype(np.uint8)
    r[:, :5] = 1
    r[:, -5:] = 1
    r[:5, :] = 1
    r[-5:, :] = 1
    return [m[1].reshape(-1), m[2].reshape(-1), r.reshape(-1)], lv[0].reshape(-1)


def mask_cases_1d(S, rng, levels):
    """As mask_cases_2d.  ComputeLevel tests the two points at exactly +-d, d =
    1, 2, .., 2^(levels-1); a point within that reach of an end must find an
    active one before it leaves the ghosts, so the band spans 2^(levels-1)
    points at each end (then the neighbour at d = 1 is active)."""
    m, lv = S.masks()
    r = (rng.random(m.shape[1]) < 0.1).astype(np.uint8)
    w = 3 + 2 ** (levels - 1)
    r[:w] = 1
    r[-w:] = 1
    return [m[1], m[2], r], lv[0]


def run_mask_algebra(A, B, cases, lv0, ncomp, q):
    """AWCM_MANAGER statics of A and B on every case: op 0 AddAdjacentZone, 1
    CheckReconstruction, 2 AddPointForDifferentiation (with ComputeLevel's
    levels of the same mask), 3 ComputeLevel (from level0), RemoveResidue on
    the scalar and the Euler field."""
    for mk in cases:
        lv_mk = B.mask_op(3, mk, lv0)[1]
        for op in range(4):
            lvl = lv_mk if op == 2 else (lv0 if op == 3 else None)
            a, b = A.mask_op(op, mk, lvl), B.mask_op(op, mk, lvl)
            assert np.array_equal(a[0], b[0]), op
            if op == 3:
                assert np.array_equal(a[1], b[1])
            if op != 3:
                assert np.all(a[0] >= mk)  # these only add points
        for nc in (1, ncomp):
            qin = q if nc == ncomp else q.reshape(-1, ncomp)[:, 0].copy()
            a, b = A.remove_residue(qin, nc, mk), B.remove_residue(qin, nc, mk)
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@needs_ref
@pytest.mark.parametrize("kind", ["quad", "dmr"])
def test_oracle_mask_algebra_matches_reference_2d(kind):
    """AWCM_MANAGER_2D's static members, oracle against the reference build, on
    masks of a 5-step run and a random banded mask."""
    L, nxc, nyc = (4, 9, 9) if kind == "quad" else (4, 17, 5)
    nx, ny = (nxc - 1) * 2 ** (L - 1) + 1, (nyc - 1) * 2 ** (L - 1) + 1
    if kind == "quad":
        q, dx, _ = O.ic2(0, nx, ny, 0.0, 1.0, 0.0, 1.0, 0.5)
        bc, xu, tf, npr = 0, 1.0, 0.25, 9
    else:
        q, dx, _ = O.ic2(1, nx, ny, 0.0, 4.0, 0.0, 1.0, 1.4)
        bc, xu, tf, npr = 1, 4.0, 0.2, 21
    S = {p: O.Solver2D(p, L, nxc, nyc, 1.4, bc, 0.0, xu, 0.0, 1.0) for p in ("orc", "ref")}
    pts = O.print_times(tf, npr)
    for s in S.values():
        s.set_q(q)
        for _ in range(5):
            s.step(1e-3, 0.4, pts)
    cases, lv0 = mask_cases_2d(S["ref"], np.random.default_rng(41))
    run_mask_algebra(S["orc"], S["ref"], cases, lv0, 4, S["ref"].get_q())


@needs_ref
def test_oracle_mask_algebra_matches_reference_1d():
    """AWCM_MANAGER_1D's static members, oracle against the reference build."""
    L = 7
    nx = 16 * 2 ** (L - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    pts = O.print_times(0.2, 41)
    S = {p: O.Solver1D(p, L, 17, 1.4, dx) for p in ("orc", "ref")}
    for s in S.values():
        s.set_q(q)
        for _ in range(20):
            s.step(1e-4, 0.5, pts)
    cases, lv0 = mask_cases_1d(S["ref"], np.random.default_rng(42), L)
    run_mask_algebra(S["orc"], S["ref"], cases, lv0, 3, S["ref"].get_q())


@needs_ref
def test_reference_breakdown_of_batch_instance_448():
    """Instance 448 of the batched workload breaks down in the reference
    scheme itself: the reference build runs 1657 steps until its time turns
    NaN (its MAX drops NaN wave speeds depending on operand order), and after
    1558 steps the next PrepareEuler leaves a non-finite wave speed, where the
    GPU stops the instance.  Both states are the committed digests, and the
    oracle reproduces the second."""
    import json
    import workloads as W
    from digest_lib import digest
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "digests_c4.json")) as f:
        d = json.load(f)
    L, rl, rr, pl, pr, u, x0, eps = W.shock_tube_instances(512)[448]
    nx = 16 * 2 ** (L - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, x0, (rl, u, pl), (rr, u, pr))
    pts = O.print_times(W.C4_T_FINAL, W.C4_NUM_PRINT)
    r = O.Solver1D("ref", L, 17, 1.4, dx)
    o = O.Solver1D("orc", L, 17, 1.4, dx)
    r.set_q(q)
    o.set_q(q)
    for _ in range(1558):
        r.step(eps, W.C4_CFL, pts)
        o.step(eps, W.C4_CFL, pts)
    o.prepare(eps)
    m, lv = o.masks()
    want = d["stop_448_1558"]
    got = digest(o.get_q(), m, lv)
    assert all(got[k] == want[k] for k in ("q", "masks", "levels")) and float(o.t[0]).hex() == want["t"]
    n = 1558
    while r.t[0] < W.C4_T_FINAL:
        r.step(eps, W.C4_CFL, pts)
        n += 1
    assert n == 1657 and np.isnan(r.t[0])
    m, lv = r.masks()
    fin = digest(r.get_q(), m, lv)
    assert all(fin[k] == d["instances"][448][k] for k in ("q", "masks", "levels"))
