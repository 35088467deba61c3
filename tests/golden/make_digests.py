"""Write tests/golden/digests_*.json from the REFERENCE build: compact digests
(SHA-256 of the bits) of state, masks and levels

* at every output time of reduced-level 2-D runs (double Mach reflection and
  the four-quadrant problem, 5 levels, over their whole schedules to t = 0.2
  and t = 0.25);
* at the first output times of the full-size runs (double Mach 4097 x 1025 to
  t = 0.05, the quadrants 2049 x 2049 to t = 0.09375);
* at the end of every instance of the 512-instance batched 1-D workload (and,
  for the instance whose scheme breaks down, where the GPU stops it).

Runs only where oracle/_ref/libawcm_ref.so exists (built by `make -C oracle`
from the reference's own headers); the files hold data only.  The full-size
runs take a few hours of host time.

    python tests/golden/make_digests.py [dmr_L5 quad_L5 dmr_full quad_full c4]
"""
import json
import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "c_pp-library_amd"))
import oracle_lib as O  # noqa: E402
from digest_lib import digest  # noqa: E402

# name: (problem, levels, coarse x, coarse y, t_final, number of output times, output times digested)
CASES_2D = {
    "dmr_L5": ("dmr", 5, 65, 17, 0.2, 21, 20),
    "quad_L5": ("quad", 5, 33, 33, 0.25, 9, 8),
    "dmr_full": ("dmr", 7, 65, 17, 0.2, 21, 5),
    "quad_full": ("quad", 7, 33, 33, 0.25, 9, 3),
}
TOL_2D, CFL_2D = 1e-3, 0.4


def run_2d(name):
    kind, L, nxc, nyc, tf, npr, upto = CASES_2D[name]
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
    out = {"problem": kind, "levels": L, "nx_coarse": nxc, "ny_coarse": nyc, "tol": TOL_2D, "cfl": CFL_2D,
           "t_final": tf, "num_print": npr, "outputs": []}
    steps = 0
    for k in range(1, upto + 1):
        while s.t[0] < pts[k]:
            s.step(TOL_2D, CFL_2D, pts)
            steps += 1
        m, lv = s.masks()
        d = digest(s.get_q(), m, lv)
        d.update(k=k, t=float(s.t[0]).hex(), steps=steps, cnt=s.cnt.value)
        out["outputs"].append(d)
        print(name, "output", k, "t", s.t[0], "steps", steps, "mask0", d["n_mask0"], flush=True)
    return out


def c4_instance(k):
    import workloads as W
    L, rl, rr, pl, pr, u, x0, eps = W.shock_tube_instances(512)[k]
    nx = 16 * 2 ** (L - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, x0, (rl, u, pl), (rr, u, pr))
    s = O.Solver1D("ref", L, 17, 1.4, dx)
    s.set_q(q)
    pts = O.print_times(W.C4_T_FINAL, W.C4_NUM_PRINT)
    steps = 0
    # the reference goes on until t >= t_final or its time is no longer finite
    # (the GPU stops an instance earlier, at its first non-finite wave speed)
    while s.t[0] < W.C4_T_FINAL and steps < 100000:
        s.step(eps, W.C4_CFL, pts)
        steps += 1
    m, lv = s.masks()
    d = digest(s.get_q(), m, lv)
    d.update(k=k, steps=steps, t=float(s.t[0]).hex())
    return d


def c4_stop_digest(k, n_steps):
    """State of instance k after n_steps steps and one more PrepareEuler: where
    the GPU stops an instance whose next wave speed is not finite."""
    import workloads as W
    L, rl, rr, pl, pr, u, x0, eps = W.shock_tube_instances(512)[k]
    nx = 16 * 2 ** (L - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, x0, (rl, u, pl), (rr, u, pr))
    s = O.Solver1D("ref", L, 17, 1.4, dx)
    s.set_q(q)
    pts = O.print_times(W.C4_T_FINAL, W.C4_NUM_PRINT)
    for _ in range(n_steps):
        s.step(eps, W.C4_CFL, pts)
    t, cnt = s.t[0], s.cnt.value
    s.prepare(eps)
    m, lv = s.masks()
    d = digest(s.get_q(), m, lv)
    d.update(k=k, steps=n_steps, t=float(t).hex(), cnt=cnt)
    return d


def run_c4():
    with mp.Pool(int(os.environ.get("DIGEST_WORKERS", "6"))) as pool:
        rows = pool.map(c4_instance, range(512), chunksize=4)
    out = {"workload": "512 seeded 1-D shock tubes (workloads.shock_tube_instances, seed 0x5EED0C4), t = 0.15",
           "instances": rows}
    for r in rows:
        if not np.isfinite(float.fromhex(r["t"])):
            out.setdefault("breakdown", []).append(r["k"])
    return out


def main():
    if not O.have_ref():
        sys.exit("oracle/_ref/libawcm_ref.so is not built (make -C oracle)")
    names = sys.argv[1:] or list(CASES_2D) + ["c4"]
    for name in names:
        if name == "c4":
            out = run_c4()
            # instance 448 breaks down: the GPU stops it after 1558 steps, at the
            # PrepareEuler whose wave speed is not finite
            out["stop_448_1558"] = c4_stop_digest(448, 1558)
        else:
            out = run_2d(name)
        with open(os.path.join(HERE, f"digests_{name}.json"), "w") as f:
            json.dump(out, f, indent=1)
        print("wrote", f"digests_{name}.json", flush=True)


if __name__ == "__main__":
    main()
