"""Write the golden vectors in tests/golden/ from the REFERENCE build.

Runs only where oracle/_ref/libawcm_ref.so exists (built by `make -C oracle`
from the reference's own headers).  Every vector below is produced by the
reference's classes driven by oracle/ref_harness.cpp; the committed .npz files
hold data only (inputs and expected outputs).

    python tests/golden/make_golden.py [file.npz ...]   (default: all files)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402


def kats(rng):
    """Per-function known-answer vectors on seeded random inputs."""
    R = O.load("ref")
    out = {}
    w = np.zeros(20)
    R.ref_weights(O.dptr(w))
    out["weights"] = w.reshape(4, 5)
    wz_in = rng.normal(size=(64, 5))
    out["wenoz_in"] = wz_in
    out["wenoz_left"] = np.array([R.ref_wenoz_left(*r) for r in wz_in])
    out["wenoz_right"] = np.array([R.ref_wenoz_right(*r) for r in wz_in])
    # 1-D face fluxes on physical states (rho>0, p>0)
    n = 64
    s1 = np.zeros((n, 6, 3))
    a1 = rng.uniform(0.5, 3.0, n)
    for k in range(n):
        rho = rng.uniform(0.1, 2.0, 6)
        u = rng.uniform(-1.5, 1.5, 6)
        p = rng.uniform(0.05, 2.0, 6)
        s1[k, :, 0] = rho
        s1[k, :, 1] = rho * u
        s1[k, :, 2] = p / 0.3999999999999999 + 0.5 * rho * u * u
    f1 = np.zeros((n, 3))
    for k in range(n):
        st = np.ascontiguousarray(s1[k]).ravel()
        R.ref_flux1(1.4, O.dptr(st), a1[k], O.dptr(f1[k]))
    out.update(flux1_states=s1, flux1_alpha=a1, flux1_out=f1)
    f1js = np.zeros((n, 3))
    for k in range(n):
        st = np.ascontiguousarray(s1[k]).ravel()
        R.ref_flux1_js(1.4, O.dptr(st), a1[k], O.dptr(f1js[k]))
    out.update(flux1_js_out=f1js)
    s2 = np.zeros((n, 6, 4))
    a2 = rng.uniform(0.5, 3.0, n)
    for k in range(n):
        rho = rng.uniform(0.1, 2.0, 6)
        u = rng.uniform(-1.5, 1.5, 6)
        v = rng.uniform(-1.5, 1.5, 6)
        p = rng.uniform(0.05, 2.0, 6)
        s2[k, :, 0] = rho
        s2[k, :, 1] = rho * u
        s2[k, :, 2] = rho * v
        s2[k, :, 3] = p / 0.3999999999999999 + 0.5 * rho * (u * u + v * v)
    f2 = np.zeros((n, 2, 4))
    for k in range(n):
        st = np.ascontiguousarray(s2[k]).ravel()
        for d in range(2):
            R.ref_flux2(1.4, d, O.dptr(st), a2[k], O.dptr(f2[k, d]))
    out.update(flux2_states=s2, flux2_alpha=a2, flux2_out=f2)
    return out


def traj1(kind, L, nxc, tol, cfl, tf, npr, snaps):
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    if kind == "sod":
        q, dx, _ = O.ic1(0, nx, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    else:
        q, dx, _ = O.ic1(1, nx, -5.0, 5.0, -4.0, O.SHU_L, O.SHU_L)
    s = O.Solver1D("ref", L, nxc, 1.4, dx)
    s.set_q(q)
    pts = O.print_times(tf, npr)
    out = dict(q0=q, dx=dx, tol=tol, cfl=cfl, levels=L, nx_coarse=nxc, print_times=pts, snaps=np.array(snaps))
    n = 0
    for target in snaps:
        while n < target:
            s.step(tol, cfl, pts)
            n += 1
        m, lv = s.masks()
        out[f"q_{target}"] = s.get_q()
        out[f"masks_{target}"] = m
        out[f"levels_{target}"] = lv
        out[f"t_{target}"] = s.t[0]
        out[f"cnt_{target}"] = s.cnt.value
    return out


def traj2(kind, L, nxc, nyc, tol, cfl, tf, npr, snaps):
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    ny = (nyc - 1) * 2 ** (L - 1) + 1
    if kind == "quad":
        q, dx, _ = O.ic2(0, nx, ny, 0.0, 1.0, 0.0, 1.0, 0.5)
        bc, xu = 0, 1.0
    else:
        q, dx, _ = O.ic2(1, nx, ny, 0.0, 4.0, 0.0, 1.0, 1.4)
        bc, xu = 1, 4.0
    s = O.Solver2D("ref", L, nxc, nyc, 1.4, bc, 0.0, xu, 0.0, 1.0)
    s.set_q(q)
    pts = O.print_times(tf, npr)
    out = dict(q0=q, dx=dx, tol=tol, cfl=cfl, levels=L, nx_coarse=nxc, ny_coarse=nyc, bc=bc, xu=xu,
               print_times=pts, snaps=np.array(snaps))
    n = 0
    for target in snaps:
        while n < target:
            s.step(tol, cfl, pts)
            n += 1
        m, lv = s.masks()
        out[f"q_{target}"] = s.get_q()
        out[f"masks_{target}"] = m.astype(np.uint8)
        out[f"levels_{target}"] = lv.astype(np.int8)
        out[f"t_{target}"] = s.t[0]
        out[f"cnt_{target}"] = s.cnt.value
    return out


def all_grid2(kind, L, nxc, nyc, steps, seed):
    """WAVELET_2D::InterpolateAllGridEuler on the state after `steps` steps,
    from mask0, from mask2, from a random mask and from no points at all (each
    from that same state), then two more steps after the one from mask0."""
    tr = traj2(kind, L, nxc, nyc, 1e-3, 0.4, 0.25 if kind == "quad" else 0.2, 9 if kind == "quad" else 21, [steps])
    bc, xu = (0, 1.0) if kind == "quad" else (1, 4.0)
    s = O.Solver2D("ref", L, nxc, nyc, 1.4, bc, 0.0, xu, 0.0, 1.0)
    q_in, m = tr[f"q_{steps}"], tr[f"masks_{steps}"]
    rng = np.random.default_rng(seed)
    froms = {"mask0": m[1], "mask2": m[3], "random": (rng.random(m[1].shape) < 0.3).astype(np.uint8),
             "none": np.zeros_like(m[1])}
    out = dict(levels=L, nx_coarse=nxc, ny_coarse=nyc, bc=bc, xu=xu, q_in=q_in, print_times=tr["print_times"],
               q0=tr["q0"], steps=steps, t_in=tr[f"t_{steps}"], cnt_in=tr[f"cnt_{steps}"])
    for name, f in froms.items():
        s.set_q(q_in)
        s.interpolate_all(f)
        out[f"from_{name}"] = f.astype(np.uint8)
        out[f"q_{name}"] = s.get_q()
    # the driver's order: its run, then the whole grid from mask0; here the run
    # then goes on for two steps from that field and mask0
    s.set_q(tr["q0"])
    for _ in range(steps):
        s.step(1e-3, 0.4, tr["print_times"])
    s.interpolate_all(s.masks()[0][1].copy())
    for _ in range(2):
        s.step(1e-3, 0.4, tr["print_times"])
    mm, lv = s.masks()
    out.update(q_after=s.get_q(), masks_after=mm.astype(np.uint8), levels_after=lv.astype(np.int8), t_after=s.t[0])
    return out


def scalar_box2(L, nxc, seed):
    """AWCM_2D::Prepare on a scalar field, three rounds, the field moved
    between them (shifted one column and damped: a stand-in for the driver's
    stages), then WAVELET_2D::Interpolate from/to four mask pairs."""
    n = (nxc - 1) * 2 ** (L - 1) + 1
    s = O.Solver2D("ref", L, nxc, nxc, 1.4, 0, 0.0, 1.0, 0.0, 1.0)
    x = (np.arange(-3, n + 3) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    q = np.exp(-50 * (X - 0.5) ** 2 - 50 * (Y - 0.3) ** 2) + 0.5 * (np.abs(X - 0.6) + np.abs(Y - 0.6) < 0.2)
    out = dict(levels=L, nx_coarse=nxc, q0=q.ravel())
    cur = q.ravel()
    for r in range(3):
        s.set_scalar(cur)
        s.prepare_scalar(1e-4)
        m, lv = s.masks()
        out[f"q_{r}"], out[f"masks_{r}"], out[f"levels_{r}"] = s.get_scalar(), m.astype(np.uint8), lv.astype(np.int8)
        cur = (np.roll(s.get_scalar().reshape(n + 6, n + 6), 1, axis=1) * 0.999).ravel()
    rng = np.random.default_rng(seed)
    m = out["masks_2"]
    pairs = {"m0_all": (m[1], np.ones_like(m[1])), "m0_m3": (m[1], m[4]), "m1_m2": (m[2], m[3]),
             "random": ((rng.random(m[1].shape) < 0.3).astype(np.uint8), (rng.random(m[1].shape) < 0.6).astype(np.uint8))}
    for name, (f, t) in pairs.items():
        s.set_scalar(out["q_2"])
        s.interpolate_scalar(f, t)
        out[f"from_{name}"], out[f"to_{name}"], out[f"qi_{name}"] = f.astype(np.uint8), t.astype(np.uint8), s.get_scalar()
    return out


def scalar_1d(L, nxc, seed):
    """AWCM_1D::Prepare on a scalar field (three rounds, the field moved
    between them) and WAVELET_1D::Interpolate from/to four mask pairs (None:
    the one-mask form, every point)."""
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    s = O.Solver1D("ref", L, nxc, 1.4, 1.0 / nx)
    x = (np.arange(-3, nx + 3) + 0.5) / nx
    q = np.exp(-200 * (x - 0.4) ** 2) + 0.5 * (np.abs(x - 0.7) < 0.1)
    out = dict(levels=L, nx_coarse=nxc, q0=q)
    cur = q
    for r in range(3):
        s.set_scalar(cur)
        s.prepare_scalar(1e-4)
        m, lv = s.masks()
        out[f"q_{r}"], out[f"masks_{r}"], out[f"levels_{r}"] = s.get_scalar(), m.astype(np.uint8), lv.astype(np.int8)
        cur = np.roll(s.get_scalar(), 2) * 0.999
    rng = np.random.default_rng(seed)
    m = out["masks_2"]
    pairs = {"m0_all": (m[1], None), "m0_m3": (m[1], m[4]), "m1_m2": (m[2], m[3]),
             "random": ((rng.random(m[1].shape) < 0.3).astype(np.uint8), (rng.random(m[1].shape) < 0.6).astype(np.uint8))}
    for name, (f, t) in pairs.items():
        s.set_scalar(out["q_2"])
        s.interpolate_scalar(f, t)
        out[f"from_{name}"], out[f"qi_{name}"] = f.astype(np.uint8), s.get_scalar()
        if t is not None:
            out[f"to_{name}"] = t.astype(np.uint8)
    return out


def euler_interp_1d(L, nxc, steps, seed):
    """WAVELET_1D::InterpolateEuler, both forms, on a Sod state after `steps`
    steps, from four mask pairs (None: the one-mask form)."""
    nx = (nxc - 1) * 2 ** (L - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, 0.5, O.SOD_L, O.SOD_R)
    s = O.Solver1D("ref", L, nxc, 1.4, dx)
    s.set_q(q)
    pts = O.print_times(0.2, 41)
    for _ in range(steps):
        s.step(1e-4, 0.5, pts)
    q_in, m = s.get_q(), s.masks()[0]
    rng = np.random.default_rng(seed)
    out = dict(levels=L, nx_coarse=nxc, dx=dx, q_in=q_in)
    pairs = {"m0_all": (m[1], None), "m0_m3": (m[1], m[4]), "m1_m2": (m[2], m[3]),
             "random_all": ((rng.random(m[1].shape) < 0.3).astype(np.uint8), None)}
    for name, (f, t) in pairs.items():
        s.set_q(q_in)
        s.interpolate_euler(f, t)
        out[f"from_{name}"], out[f"q_{name}"] = f.astype(np.uint8), s.get_q()
        if t is not None:
            out[f"to_{name}"] = t.astype(np.uint8)
    return out


def main():
    if not O.have_ref():
        sys.exit("oracle/_ref/libawcm_ref.so is not built (make -C oracle)")
    files = {
        "kat_functions.npz": lambda: kats(np.random.default_rng(20261203)),
        "sod_1d_L6.npz": lambda: traj1("sod", 6, 9, 1e-4, 0.5, 0.2, 41, [1, 5, 40]),
        "shu_osher_1d_L6.npz": lambda: traj1("shu", 6, 9, 1e-4, 0.5, 1.8, 41, [1, 10, 30]),
        "quad_2d_L4.npz": lambda: traj2("quad", 4, 9, 9, 1e-3, 0.4, 0.25, 9, [1, 6]),
        "dmr_2d_L4.npz": lambda: traj2("dmr", 4, 17, 6, 1e-3, 0.4, 0.2, 21, [1, 6]),
        "all_grid_quad_2d_L4.npz": lambda: all_grid2("quad", 4, 5, 5, 6, 31),
        "all_grid_dmr_2d_L3.npz": lambda: all_grid2("dmr", 3, 17, 5, 6, 32),
        "scalar_box_2d_L4.npz": lambda: scalar_box2(4, 5, 33),
        "scalar_1d_L7.npz": lambda: scalar_1d(7, 17, 34),
        "euler_interp_1d_L6.npz": lambda: euler_interp_1d(6, 17, 20, 35),
    }
    for name in sys.argv[1:] or list(files):
        np.savez_compressed(os.path.join(HERE, name), **files[name]())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
