"""GPU discrete adjoint of the 1-D AWCM step (frozen masks, levels, dt, alphas).

What is checked, through the C ABI:
* recording does not change the forward (bit-identical to awcm_run);
* the frozen replay reproduces the recorded run bit for bit, and from a
  perturbed state is bit-identical to the oracle's frozen replay -- so finite
  differences of either are finite differences of the same map;
* the adjoint is deterministic and independent of the recompute segment, and
  bit-identical to the host emulation of the same device code;
* the adjoint against central finite differences of the frozen map, on seeded
  smooth random density directions: on a short horizon to 1e-6 relative; on
  the full Shu-Osher run (4097 points, t = 1.8, 7090 steps) within the finite
  differences' own resolution, measured as the spread between steps h = 1e-8
  and h = 1e-9 (WENO-Z's |b0 - b2| kinks make the frozen map only piecewise
  smooth, so larger steps measure secants across kinks; see FD_H below);

Why density directions: in the Shu-Osher state the left region is uniform and
the right region has u = 0, p = 1 exactly, where the two acoustic
characteristic variables of the Roe projection are uniform (H * 0.4/c^2 = 1).
On a uniform WENO-Z stencil the weights are homogeneous of degree 0 in the
stencil differences, so along momentum or energy perturbations there the map
is only directionally differentiable (finite differences do not converge to
any linear map).  Density perturbations at fixed momentum and energy keep
u = 0 and p = 1, excite only the entropy wave, and are directions along which
the frozen map is differentiable.
* the same gradient against finite differences of the CPU oracle's replay;
* a torch.autograd.Function wrapper (test harness only) around replay/adjoint.
"""
import ctypes as C
import time

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

SHU_AMP = 0.2
SHORT_T = 0.04  # about 40 steps at 1025 points


def _pyawcm():
    import pyawcm
    pyawcm.lib()
    return pyawcm


def shu_ic(L, amp=SHU_AMP):
    nx = 16 * 2 ** (L - 1) + 1
    q, dx, x0 = O.ic1(1, nx, -5.0, 5.0, -4.0, O.SHU_L, O.SHU_L, amp=amp)
    return q, dx, x0, nx


def target_density(P, L, t_final, tol, amp):
    """Final density of the plain forward run from a perturbed right-state amplitude."""
    q, dx, _, nx = shu_ic(L, amp)
    s = P.Solver(1, L, 17, 1.4, dx)
    s.set_schedule(O.print_times(t_final, 2))
    s.upload_q(q)
    s.run(tol, 0.5, t_final, 1 << 20)
    return s.download_q().reshape(-1, 3)[3:-3, 0].copy()


def objective(qf, rho_t):
    r = qf.reshape(-1, 3)[3:-3, 0] - rho_t
    return float(np.dot(r, r))


def objective_grad(qf, rho_t):
    g = np.zeros_like(qf).reshape(-1, 3)
    g[3:-3, 0] = 2.0 * (qf.reshape(-1, 3)[3:-3, 0] - rho_t)
    return g.reshape(-1)


def smooth_directions(nx, x0, dx, n, seed, lo=-3.9, modes=16):
    """n density directions: N(0,1) coefficients on sin(k pi s), k = 1..modes,
    s = (x - lo)/(5 - lo) clipped to [0, 1]; momentum and energy unperturbed."""
    rng = np.random.default_rng(seed)
    x = x0 + dx * np.arange(nx)
    s = np.clip((x - lo) / (5.0 - lo), 0.0, 1.0)
    basis = np.stack([np.sin((k + 1) * np.pi * s) for k in range(modes)])
    out = []
    for _ in range(n):
        v = np.zeros((nx + 6, 3))
        v[3:-3, 0] = rng.standard_normal(modes) @ basis
        out.append(v.reshape(-1))
    return out


class Frozen:
    """A recorded GPU trajectory and its frozen map J(q0)."""

    def __init__(self, P, L, t_final, tol, ckpt, rho_t):
        self.q0, dx, self.x0, self.nx = shu_ic(L)
        self.dx = dx
        self.s = P.Solver(1, L, 17, 1.4, dx)
        self.s.set_schedule(O.print_times(t_final, 2))
        self.s.upload_q(self.q0)
        t0 = time.perf_counter()
        self.steps = self.s.adjoint_record(tol, 0.5, t_final, 20000, ckpt)
        self.record_s = time.perf_counter() - t0
        self.qf = self.s.download_q()
        self.rho_t = rho_t

    def J(self, q0):
        self.s.upload_q(q0)
        self.s.adjoint_replay()
        return objective(self.s.download_q(), self.rho_t)

    def fd(self, v, h):
        return (self.J(self.q0 + h * v) - self.J(self.q0 - h * v)) / (2 * h)

    def grad(self, ckpt=0):
        return self.s.adjoint(objective_grad(self.qf, self.rho_t), ckpt)


@pytest.fixture(scope="module")
def short():
    P = _pyawcm()
    L, tol = 7, 1e-4
    rng = np.random.default_rng(11)
    rho_t = target_density(P, L, SHORT_T, tol, SHU_AMP * (1 + 0.05 * rng.standard_normal()))
    return Frozen(P, L, SHORT_T, tol, 8, rho_t)


def test_record_is_the_plain_forward(short):
    P = _pyawcm()
    s = P.Solver(1, 7, 17, 1.4, short.dx)
    s.set_schedule(O.print_times(SHORT_T, 2))
    s.upload_q(short.q0)
    steps, _ = s.run(1e-4, 0.5, SHORT_T, 1 << 20)
    assert steps == short.steps and short.steps > 30
    assert np.array_equal(s.download_q().view(np.uint64), short.qf.view(np.uint64))


def test_replay_reproduces_recording(short):
    short.s.upload_q(short.q0)
    short.s.adjoint_replay()
    assert np.array_equal(short.s.download_q().view(np.uint64), short.qf.view(np.uint64))


def test_replay_is_the_oracle_frozen_map(short):
    """Same tape on both sides (the oracle records its own run, which is bit-
    identical to the GPU's); from a perturbed state both replays agree bit for bit."""
    o = O.Solver1D("orc", 7, 17, 1.4, short.dx)
    o.set_q(short.q0)
    o.record()
    pts = O.print_times(SHORT_T, 2)
    while o.t[0] < SHORT_T:
        o.step(1e-4, 0.5, pts)
    assert o.tape_steps() == short.steps
    v = smooth_directions(short.nx, short.x0, short.dx, 1, seed=3)[0]
    qp = short.q0 + 1e-6 * v
    o.set_q(qp)
    o.replay()
    short.s.upload_q(qp)
    short.s.adjoint_replay()
    assert np.array_equal(short.s.download_q().view(np.uint64), o.get_q().view(np.uint64))


def test_adjoint_deterministic_and_segment_independent(short):
    g8 = short.grad(8)
    print("adjoint gradient: max |g| %.3e" % np.abs(g8).max(), flush=True)
    assert np.all(np.isfinite(g8))
    assert np.array_equal(g8.view(np.uint64), short.grad(8).view(np.uint64))
    assert np.array_equal(g8.view(np.uint64), short.grad(16).view(np.uint64))
    assert np.array_equal(g8.view(np.uint64), short.grad(64).view(np.uint64))
    assert np.all(g8.reshape(-1, 3)[:3] == 0) and np.all(g8.reshape(-1, 3)[-3:] == 0)
    assert 0 < np.abs(g8).max() < 1e3
    # one finite-difference direction already pins the scale (a wrong sweep
    # that is merely deterministic fails here)
    v = smooth_directions(short.nx, short.x0, short.dx, 1, seed=5)[0]
    fd = short.fd(v, 1e-6)
    assert abs(float(np.dot(g8, v)) - fd) <= 1e-6 * abs(fd), (float(np.dot(g8, v)), fd)


def test_adjoint_matches_host_emulation(short, tmp_path):
    """The same device code compiled for the host (tests/native/emu1d.cpp)
    records the same trajectory and gives the same gradient: identical, or
    within a few ulp where a device FP64 division rounds a halfway case the
    other way."""
    import emu_lib
    E = emu_lib.build(str(tmp_path / "emu1d.so"))
    _dp = C.POINTER(C.c_double)
    pts = O.print_times(SHORT_T, 2)
    h = E.emu_create(7, 17, 1.4, short.dx, pts.ctypes.data_as(_dp), len(pts))
    try:
        q0 = short.q0.copy()
        E.emu_set_q(h, q0.ctypes.data_as(_dp))
        assert E.emu_record(h, 1e-4, 0.5, SHORT_T, 20000, 8) == short.steps
        gf = objective_grad(short.qf, short.rho_t)
        ge = np.zeros_like(gf)
        E.emu_adjoint(h, gf.ctypes.data_as(_dp), 0, ge.ctypes.data_as(_dp))
    finally:
        E.emu_destroy(h)
    g = short.grad()
    same = np.array_equal(g.view(np.uint64), ge.view(np.uint64))
    err = np.abs(g - ge).max() / np.abs(ge).max()
    print("GPU vs host-emulated adjoint: bit-identical %s, max rel diff %.3e" % (same, err))
    assert err <= 1e-13


def test_adjoint_matches_fd_short_horizon(short):
    g = short.grad()
    errs = []
    for v in smooth_directions(short.nx, short.x0, short.dx, 16, seed=2026):
        ad = float(np.dot(g, v))
        fd = short.fd(v, 1e-6)
        errs.append(abs(ad - fd) / abs(fd))
    print("short-horizon adjoint vs FD, max rel err", max(errs))
    assert max(errs) <= 1e-6, errs


def test_torch_autograd_wrapper(short):
    torch = pytest.importorskip("torch")

    class FrozenStep(torch.autograd.Function):
        """Frozen AWCM trajectory as an autograd op (test harness only)."""

        @staticmethod
        def forward(ctx, q0):
            short.s.upload_q(q0.detach().cpu().numpy())
            short.s.adjoint_replay()
            return torch.from_numpy(short.s.download_q())

        @staticmethod
        def backward(ctx, gout):
            return torch.from_numpy(short.s.adjoint(gout.detach().cpu().numpy().copy()))

    q0 = torch.from_numpy(short.q0.copy()).requires_grad_(True)
    qf = FrozenStep.apply(q0)
    rho_t = torch.from_numpy(short.rho_t)
    J = ((qf.view(-1, 3)[3:-3, 0] - rho_t) ** 2).sum()
    J.backward()
    v = smooth_directions(short.nx, short.x0, short.dx, 1, seed=77)[0]
    ad = float(torch.dot(q0.grad, torch.from_numpy(v)))
    fd = short.fd(v, 1e-6)
    assert abs(ad - fd) <= 1e-6 * abs(fd)
    # torch's sum and numpy's dot add the squares in different orders
    Jn = objective(short.qf, short.rho_t)
    assert abs(float(J.detach()) - Jn) <= 1e-12 * abs(Jn)


@pytest.fixture(scope="module")
def shu_osher():
    """The full Shu-Osher adjoint case: 9 levels (4097 points), eps 1e-4, CFL 0.5, t = 1.8."""
    P = _pyawcm()
    L, tol = 9, 1e-4
    rng = np.random.default_rng(2026)
    delta = 0.05 * rng.standard_normal()
    t0 = time.perf_counter()
    rho_t = target_density(P, L, 1.8, tol, SHU_AMP * (1 + delta))
    fwd_s = time.perf_counter() - t0
    F = Frozen(P, L, 1.8, tol, 32, rho_t)
    F.fwd_s = fwd_s
    print(f"Shu-Osher: {F.steps} steps, forward {fwd_s:.3f} s, record {F.record_s:.3f} s")
    return F


def backward(F):
    t0 = time.perf_counter()
    F.g = F.grad()
    F.adj_s = time.perf_counter() - t0


def test_shu_osher_record_and_replay(shu_osher):
    F = shu_osher
    assert F.steps > 5000
    F.s.upload_q(F.q0)
    F.s.adjoint_replay()
    assert np.array_equal(F.s.download_q().view(np.uint64), F.qf.view(np.uint64))


@pytest.fixture(scope="module")
def shu_oracle(shu_osher):
    """The CPU oracle's recording of the same run (bit-identical tape) and its
    frozen map J(q0)."""
    F = shu_osher
    o = O.Solver1D("orc", 9, 17, 1.4, F.dx)
    o.set_q(F.q0)
    o.record()
    pts = O.print_times(1.8, 2)
    while o.t[0] < 1.8:
        o.step(1e-4, 0.5, pts)
    assert o.tape_steps() == F.steps

    def J(q0):
        o.set_q(q0)
        o.replay()
        return objective(o.get_q(), F.rho_t)

    return J


def test_shu_osher_gpu_frozen_map_is_the_oracle_one(shu_osher, shu_oracle):
    """Finite differences of the GPU replay and of the oracle's frozen replay
    (the CPU reference of the gradient) are the same numbers."""
    F, J = shu_osher, shu_oracle
    v = smooth_directions(F.nx, F.x0, F.dx, 1, seed=99)[0]
    fd_cpu = (J(F.q0 + 1e-6 * v) - J(F.q0 - 1e-6 * v)) / 2e-6
    assert fd_cpu == F.fd(v, 1e-6)


# Over the full horizon the frozen map is only piecewise smooth (the |b0 - b2|
# kinks of WENO-Z are crossed by a perturbation of size h somewhere in 7090
# steps x 4097 points x 3 stages), so a central difference at h = 1e-6 is a
# secant across kinks (off by up to ~3e-2 from the derivative at q0), and
# differences converge to the adjoint only as h -> 0: measured on this case,
# the relative gap is ~1e-2 at h = 1e-6, ~1e-4..5e-3 at 1e-7, ~3e-6..4e-4 at
# 1e-8 and ~2e-6..1e-4 at 1e-9.  The adjoint is therefore held to the finite
# differences' own resolution: |ad - FD(1e-9)| <= 4 |FD(1e-8) - FD(1e-9)| +
# 1e-6 |FD(1e-9)|.
FD_H = (1e-8, 1e-9)


def _fd_pair(F, v):
    return F.fd(v, FD_H[0]), F.fd(v, FD_H[1])


def test_shu_osher_adjoint_vs_fd(shu_osher):
    F = shu_osher
    backward(F)
    assert F.steps > 5000
    rows = []
    for v in smooth_directions(F.nx, F.x0, F.dx, 16, seed=1234):
        ad = float(np.dot(F.g, v))
        f8, f9 = _fd_pair(F, v)
        rows.append((ad, f8, f9))
    errs = [abs(a - f9) / abs(f9) for a, _, f9 in rows]
    spread = [abs(f8 - f9) / abs(f9) for _, f8, f9 in rows]
    print("full-horizon adjoint vs FD(1e-9) on 16 directions: max rel gap %.3e, median %.3e; "
          "FD(1e-8) vs FD(1e-9): max %.3e" % (max(errs), float(np.median(errs)), max(spread)))
    for a, f8, f9 in rows:
        assert abs(a - f9) <= 4 * abs(f8 - f9) + 1e-6 * abs(f9), (a, f8, f9)
    assert float(np.median(errs)) <= 1e-4


def test_shu_osher_adjoint_vs_oracle_fd(shu_osher, shu_oracle):
    """Finite differences of the CPU oracle's frozen replay are those of the
    GPU replay, bit for bit, and the adjoint meets them as above."""
    F, J = shu_osher, shu_oracle
    backward(F)
    v = smooth_directions(F.nx, F.x0, F.dx, 1, seed=99)[0]
    h = FD_H[1]
    fd_cpu = (J(F.q0 + h * v) - J(F.q0 - h * v)) / (2 * h)
    f8, f9 = _fd_pair(F, v)
    assert fd_cpu == f9  # same frozen map, bit-identical replays
    ad = float(np.dot(F.g, v))
    assert abs(ad - fd_cpu) <= 4 * abs(f8 - fd_cpu) + 1e-6 * abs(fd_cpu), (ad, f8, fd_cpu)


def test_shu_osher_adjoint_cost(shu_osher):
    """Adjoint (recompute from checkpoints + reverse sweep) against the plain forward run."""
    F = shu_osher
    backward(F)
    assert F.adj_s <= 4.0 * F.fwd_s, (F.adj_s, F.fwd_s)
