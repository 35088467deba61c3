#!/usr/bin/env python3
"""Headline benchmark: 2-D double Mach reflection AWCM forward pass on one MI355X.

Workload (the configuration the headline metric is quoted on): Mach-10 shock
at 60 degrees on [0,4]x[0,1], finest grid 4097 x 1025 (7 levels, 65 x 17
coarse), pre-shock state (1.4, 0, 0, 1), eps = 1e-3, WENO-Z5 + Roe, SSP-RK3 at
CFL 0.4, output schedule of 21 times to t = 0.2, FP64.

A step is one full AWCM time step (PrepareEuler + CFL dt + three SSP-RK3
stages + post-processing) over the device-resident state, replayed as one HIP
graph.  The run:

1. the whole DMR run from t = 0 to t = 0.2 on the GPU, timed with HIP events
   on the solver's stream in chunks of steps (full_run: updates/s over the
   run), with the state snapshotted as it passes t = 0.02, 0.1 and 0.19;
2. from each snapshot, --warmup untimed steps, then exactly --steps timed
   steps (windows); the t ~ 0.1 window is `value`.  Each window is replayed
   twice more from the same state, bit-identically: with HIP events around
   every kernel class (breakdown, roofline of k2_stage_faces) and step by step
   with the masks read back (the algorithmic HBM bytes of each step);
3. the four-quadrant problem at full size (2049 x 2049), a window at its first
   output time;
4. the batched 1-D form, the adjoint and the scalar box-advection driver's
   loop (main_2d_advect.cpp) on the device;
5. the reference's CPU path (oracle/_ref, the reference's own headers) on a
   bounded sample of the headline window, with all host threads and with one,
   and on every 8th instance of the batched workload (one host process per
   core, one instance at a time each).

value = active-point RK3 updates per second = sum over the timed steps of
|mask0| / device time.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--only T] [--no-full] [--no-cpu] [--no-batch] [--no-c2]
    python bench.py --hbm-copy     # a plain device copy of 1 GiB (HBM counter calibration)
    python bench.py --only 0.1 --save-snapshot s.npz   # then, e.g. under rocprofv3:
    python bench.py --load-snapshot s.npz --steps 5 --warmup 1 --no-cpu --no-batch --no-c2
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "c_pp-library_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "active-point RK3 updates/s & %HBM peak; batched vs plain; adjoint/fwd cost"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# DMR configuration
LEVELS, NXC, NYC = 7, 65, 17
RES = (NYC - 1) * 2 ** (LEVELS - 1) + 1  # 1025
TOL, CFL, T_FINAL, NUM_PRINT, RHO_PRE = 1e-3, 0.4, 0.2, 21, 1.4
WINDOWS = (0.02, 0.1, 0.19)
HEADLINE = 0.1
# four-quadrant configuration (full size)
Q_LEVELS, Q_NXC = 7, 33
Q_RES = (Q_NXC - 1) * 2 ** (Q_LEVELS - 1) + 1  # 2049
Q_TOL, Q_CFL, Q_T_FINAL, Q_NUM_PRINT = 1e-3, 0.4, 0.25, 9

# Algorithmic cost of the dominant kernel, k2_stage_faces (one WENO-Z/Roe face
# flux per thread, each face of the stage once).  FP64 operations of one face
# flux as the reference writes it (+ - * / sqrt, WENOZ.h:236-320 with
# ROESOLVER.h / EULER_EQUATION.h), counted mechanically: 1202.
FLOPS_PER_FACE = 1202
# Bytes per face: the face's own state (32 B, neighbours' states are the
# neighbouring faces' own reads), its 4 B list entry, and the 32 B flux written.
BYTES_PER_FACE = 32 + 4 + 32
# Measured with rocprofv3 on the t ~ 0.1 window (profiles/
# k2_stage_faces_dmr_t0.1_pmc.md): HBM bytes per face of a k2_stage_faces
# launch, FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors from `bench.py
# --hbm-copy`), and issued FP64 lane operations per face.
TRAFFIC_BYTES_PER_FACE = 183
ISSUED_FP64_OPS_PER_FACE = 1803


def ensure_built():
    import pyawcm
    if not os.path.exists(pyawcm.LIB_PATH):
        subprocess.run([sys.executable, os.path.join(ROOT, "c_pp-library_amd", "build.py"), "-q"], check=True)
    return pyawcm


def cpu_sample(path, budget_s):
    """Time DMR steps of the reference build (or the oracle port) on the host
    from the state saved in `path`; OpenMP threads from OMP_NUM_THREADS."""
    import oracle_lib as O
    z = np.load(path)
    kind, prefix = ("reference", "ref") if O.have_ref() else ("port", "orc")
    s = O.Solver2D(prefix, LEVELS, NXC, NYC, 1.4, 1, 0.0, 4.0, 0.0, 1.0)
    s.set_q(z["q"])
    s.set_mask0(z["m0"])
    s.t[0] = float(z["t"])
    s.cnt.value = int(z["cnt"])
    pts = O.print_times(T_FINAL, NUM_PRINT)
    upd, elapsed, steps = 0, 0.0, 0
    while elapsed < budget_s and steps < 50:
        t0 = time.perf_counter()
        s.step(TOL, CFL, pts)
        elapsed += time.perf_counter() - t0
        steps += 1
        m, _ = s.masks()
        upd += int(m[1, 3:-3, 3:-3].sum())
    return {"value": upd / elapsed, "kind": kind, "steps": steps, "elapsed_s": elapsed, "prefix": prefix}


def cpu_baseline(snap, budget_s, threads):
    """The CPU sample in a child process with `threads` OpenMP threads."""
    q, m0, t, cnt = snap
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "state.npz")
        np.savez(path, q=q, m0=m0, t=t, cnt=cnt)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-sample", path, "--cpu-budget",
                            str(budget_s)], capture_output=True, text=True, env=env, timeout=budget_s * 6 + 600)
    if r.returncode != 0:
        return {"error": r.stderr[-400:]}
    res = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": res["value"], "unit": "updates/s", "cores": threads, "kind": res["kind"],
            "sample": f"{res['steps']} DMR steps of the headline window, resumed from the GPU state at t={t:.6f} "
                      f"({res['elapsed_s']:.1f} s wall, OpenMP {threads} thread(s), {res['prefix']} build)"}


def c4_cpu_instance(k):
    """Instance k of the batched workload run to t = 0.15 on the host by the
    reference build (or the oracle port), the way make_digests.py runs it."""
    import oracle_lib as O
    import workloads as W
    L, rl, rr, pl, pr, u, x0, eps = W.shock_tube_instances(512)[k]
    nx = 16 * 2 ** (L - 1) + 1
    q, dx, _ = O.ic1(0, nx, 0.0, 1.0, x0, (rl, u, pl), (rr, u, pr))
    s = O.Solver1D("ref" if O.have_ref() else "orc", L, 17, 1.4, dx)
    s.set_q(q)
    pts = O.print_times(W.C4_T_FINAL, W.C4_NUM_PRINT)
    steps = upd = 0
    while s.t[0] < W.C4_T_FINAL and np.isfinite(s.t[0]) and steps < 100000:
        s.step(eps, W.C4_CFL, pts)
        steps += 1
        m, _ = s.masks()
        upd += int(m[1, 3:-3].sum())
    return k, steps, upd


def c4_cpu_sample(stride, workers):
    """The sampled instances (every stride-th, from instance 3) on `workers`
    host processes, one instance at a time each; wall time of the whole pool."""
    import multiprocessing as mp
    import oracle_lib as O
    ks = list(range(3, 512, stride))
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        res = pool.map(c4_cpu_instance, ks, chunksize=1)
    wall = time.perf_counter() - t0
    return {"instances": len(ks), "stride": stride, "wall_s": wall, "steps": sum(r[1] for r in res),
            "updates": sum(r[2] for r in res), "kind": "reference" if O.have_ref() else "port"}


def c4_cpu_baseline(workers, stride=8):
    """c4_cpu_sample in a child process that never touches the GPU (one OpenMP
    thread per worker process)."""
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--c4-cpu-sample", str(stride),
                        "--c4-cpu-workers", str(workers)], capture_output=True, text=True, env=env, timeout=1200)
    if r.returncode != 0:
        return {"error": r.stderr[-400:]}
    res = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": res["updates"] / res["wall_s"], "unit": "updates/s", "cores": workers, "kind": res["kind"],
            "sample": f"{res['instances']} of the 512 instances (every {stride}th from instance 3) run to t = 0.15 "
                      f"on {workers} host processes, one instance each at a time: {res['steps']} steps, "
                      f"{res['updates']} updates in {res['wall_s']:.1f} s wall"}


def hbm_copy():
    """A plain device copy of known size, for scaling the FETCH_SIZE/WRITE_SIZE
    counters: 1 GiB filled, then copied device to device, three times."""
    import torch
    n = (1 << 30) // 8
    a = torch.empty(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"hbm_copy_bytes": n * 8, "copies": 3}))


def batched_section(P, cmp_steps, cpu_workers=0):
    """512 seeded 1-D shock tubes (uneven 513..4097-point grids, eps 1e-3/1e-4,
    CFL 0.5, t = 0.15) in one graph-replayed persistent launch with the
    autotuned shape, the same work as one plain launch per instance, and (with
    cpu_workers) a sample of the instances on the host's cores."""
    import workloads as W
    b, inst, params, _ = W.shock_tube_batch(512)
    threads, wgs, probe = b.autotune(10)
    t0 = time.perf_counter()
    b.run()
    full_s = time.perf_counter() - t0
    states = [b.state(k) for k in range(512)]
    updates = sum(s["updates"] for s in states)
    b.close()
    # batched vs sequential on identical work: every instance advances cmp_steps
    # steps from its initial state (results are bit-identical between the two)
    bb, _, _, _ = W.shock_tube_batch(512)
    bb.set_launch(threads, wgs)
    t0 = time.perf_counter()
    bb.run(cmp_steps)
    bat_s = time.perf_counter() - t0
    bb.close()
    bs, _, _, _ = W.shock_tube_batch(512)
    t0 = time.perf_counter()
    bs.run_sequential(cmp_steps)
    seq_s = time.perf_counter() - t0
    bs.close()
    out = {
        "workload": "512 seeded 1-D shock tubes (splitmix64 seed 0x5EED0C4), 513..4097 finest points, "
                    "eps 1e-3/1e-4, WENO-Z5+Roe, SSP-RK3 CFL 0.5 to t=0.15, FP64",
        "launch": {"threads": threads, "workgroups": wgs or 512, "probe_ms": [round(x, 3) for x in probe]},
        "full_run_s": full_s,
        "updates_per_s": updates / full_s,
        "steps_total": sum(s["steps"] for s in states),
        "instances_reaching_t_final": sum(1 for s in states if s["status"] == 0),
        "instances_nonfinite": sum(1 for s in states if s["status"] == -4),
        "vs_sequential": {"steps_per_instance": cmp_steps, "batched_s": bat_s, "sequential_s": seq_s,
                          "speedup": seq_s / bat_s},
    }
    if cpu_workers:
        cb = c4_cpu_baseline(cpu_workers)
        out["cpu_baseline"] = cb
        if "value" in cb and cb["value"] > 0:
            out["gpu_over_cpu"] = out["updates_per_s"] / cb["value"]
    return out


def adjoint_section(P):
    """The Shu-Osher adjoint case on the GPU (4097 points, eps 1e-4, CFL 0.5,
    t = 1.8): plain forward run, tape recording (the forward with its tape),
    frozen replay, and the gradient of J = sum (rho_final - rho_target)^2 by
    checkpoint-and-recompute every 32 steps.  Cost ratios are against the
    plain forward run."""
    q, dx, _ = P.ic_shu_osher_1d(4097)
    pts = P.print_times(1.8, 2)
    s = P.Solver(1, 9, 17, 1.4, dx)
    s.set_schedule(pts)
    s.upload_q(q)
    t0 = time.perf_counter()
    n, _ = s.run(1e-4, 0.5, 1.8, 1 << 20)
    fwd = time.perf_counter() - t0
    qt, _, _ = P.ic_shu_osher_1d(4097, amp=0.2 * 1.05)
    st = P.Solver(1, 9, 17, 1.4, dx)
    st.set_schedule(pts)
    st.upload_q(qt)
    st.run(1e-4, 0.5, 1.8, 1 << 20)
    rho_t = st.download_q().reshape(-1, 3)[3:-3, 0].copy()
    s2 = P.Solver(1, 9, 17, 1.4, dx)
    s2.set_schedule(pts)
    s2.upload_q(q)
    t0 = time.perf_counter()
    n2 = s2.adjoint_record(1e-4, 0.5, 1.8, 20000, 32)
    rec = time.perf_counter() - t0
    qf = s2.download_q()
    s2.upload_q(q)
    t0 = time.perf_counter()
    s2.adjoint_replay()
    rep = time.perf_counter() - t0
    gf = np.zeros((qf.size // 3, 3))
    gf[3:-3, 0] = 2.0 * (qf.reshape(-1, 3)[3:-3, 0] - rho_t)
    t0 = time.perf_counter()
    g = s2.adjoint(gf.reshape(-1), 32)
    bwd = time.perf_counter() - t0
    return {"workload": "Shu-Osher 4097 points (9 levels), eps 1e-4, CFL 0.5, t=1.8; J = sum (rho_T - rho_target)^2, "
                        "target from the right-state amplitude 0.2*1.05; checkpoints every 32 steps",
            "steps": n, "forward_s": fwd, "record_s": rec, "replay_s": rep, "steps_recorded": n2,
            "backward_s": bwd, "gradient_finite": bool(np.all(np.isfinite(g))),
            "adjoint_over_forward": bwd / fwd, "record_plus_backward_over_forward": (rec + bwd) / fwd}


def dmr_solver(P):
    q0, dx = P.ic_double_mach_2d(RES, RHO_PRE)
    S = P.Solver(2, LEVELS, NXC, 1.4, dx, ny_coarse=NYC, x0=0.0, bc_kind=P.BC_DOUBLE_MACH)
    S.set_schedule(P.print_times(T_FINAL, NUM_PRINT))
    S.upload_q(q0)
    return S


def snapshot(S):
    S.synchronize()
    return (S.download_q(), S.masks()[0][1].copy(), S.t, S.cnt)


def restore(S, snap):
    q, m0, t, cnt = snap
    S.upload_q(q)
    S.set_mask0(m0)
    S.t, S.cnt = t, cnt


def full_run(S, t_final, marks, chunk=32):
    """The run to t_final in timed chunks of graph-replayed steps, with a
    snapshot as t passes each mark (between chunks, outside the timed spans)."""
    ms = upd = steps = 0
    snaps = {}
    while S.t < t_final:
        for w in marks:
            if w not in snaps and S.t >= w:
                snaps[w] = snapshot(S)
        n, u, dt_ms = S.run_timed(TOL, CFL, t_final, chunk)
        ms, upd, steps = ms + dt_ms, upd + u, steps + n
        if n == 0:
            break
    return {"steps": steps, "updates": upd, "device_s": ms / 1e3, "updates_per_s": upd / (ms / 1e3),
            "ms_per_step": ms / max(steps, 1), "t_end": S.t}, snaps


def step_bytes(m, prev_m0):
    """Algorithmic HBM bytes of one 2-D step from its masks (interior points):
    s(2|M2| + 3|M1| + 3|M0| + |M3| - |M0|) RK stages and stencil halos,
    16(|P| + |M3|) density transforms, 8(nc-1)|M3 minus P| interpolation,
    5(|M0|+|M1|+|M2|+|M3|) list entries and levels; s = 32 B, nc = 4."""
    M = [int(m[k, 3:-3, 3:-3].sum()) for k in range(1, 5)]
    P0 = int(prev_m0[3:-3, 3:-3].sum())
    new = int((m[4, 3:-3, 3:-3] & (1 - prev_m0[3:-3, 3:-3])).sum())
    s = 32
    return s * (2 * M[2] + 3 * M[1] + 3 * M[0] + M[3] - M[0]) + 16 * (P0 + M[3]) + 8 * 3 * new + 5 * sum(M)


def window(P, S, snap, K, W, detail=True):
    """W warm-up steps from snap, then K timed steps (graph replay); with
    detail, the same K steps again with per-kernel events and step by step with
    the masks read back (both bit-identical to the timed replay)."""
    restore(S, snap)
    S.run(TOL, CFL, T_FINAL, W)
    start = snapshot(S)
    steps, upd, ms = S.run_timed(TOL, CFL, T_FINAL, K)
    t_end = S.t
    res = {"t_start": start[2], "t_end": t_end, "steps": steps, "updates": upd, "device_ms": ms,
           "ms_per_step": ms / max(steps, 1), "updates_per_s": upd / (ms / 1e3),
           "mean_active_points": upd / max(steps, 1)}
    if not detail:
        return res, start
    restore(S, start)
    S.profile(True)
    steps_p, upd_p, ms_p = S.run_timed(TOL, CFL, T_FINAL, K)
    w_ms, w_launch, w_pts = S.profile_read(P.KERNEL_STAGE_WENO)
    c_ms, c_launch, c_pts = S.profile_read(P.KERNEL_STAGE_COARSE)
    p_ms, p_launch, _ = S.profile_read(P.KERNEL_PREPARE)
    S.profile(False)
    assert steps_p == steps and upd_p == upd and S.t == t_end, "profiled replay diverged"
    res["eager_ms_per_step"] = ms_p / max(steps, 1)
    res["breakdown_ms_per_step"] = {"stage_faces": w_ms / max(steps, 1), "stage_update": c_ms / max(steps, 1),
                                    "prepare_euler": p_ms / max(steps, 1),
                                    "other": (ms_p - w_ms - c_ms - p_ms) / max(steps, 1)}
    res["faces"] = {"launches": w_launch, "faces": w_pts, "ms": w_ms}
    # algorithmic bytes, step by step
    restore(S, start)
    prev = S.masks()[0][1].copy()
    nbytes = 0
    for _ in range(K):
        S.run(TOL, CFL, T_FINAL, 1)
        m, _ = S.masks()
        nbytes += step_bytes(m, prev)
        prev = m[1].copy()
    assert S.t == t_end, "step-by-step replay diverged"
    gbs = nbytes / (ms / 1e3) / 1e9
    res["algorithmic_bytes_per_step"] = nbytes / max(steps, 1)
    res["algorithmic_gbs"] = gbs
    res["hbm_frac"] = gbs / HBM_PEAK_GBS
    return res, start


def faces_roofline(win):
    f = win["faces"]
    avg_launch_s = (f["ms"] / f["launches"]) / 1e3
    per_launch = f["faces"] / f["launches"]
    achieved = per_launch * FLOPS_PER_FACE / avg_launch_s / 1e12
    gbs = per_launch * BYTES_PER_FACE / avg_launch_s / 1e9
    traffic = per_launch * TRAFFIC_BYTES_PER_FACE if TRAFFIC_BYTES_PER_FACE else None
    return {
        "bound": "fp64-valu",
        "compute": "FP64 vector ALU: k2_stage_faces issues no MFMA; MI355X's dense FP64 peak is 78.6 TFLOP/s, and "
                   "bit-exact parity forbids FMA contraction",
        "kernel": "k2_stage_faces (WENO-Z/Roe face fluxes, each face of a stage once), headline window",
        "achieved": achieved,
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / FP64_PEAK_TFLOPS,
        "traffic": traffic,
        "traffic_over_algorithmic": traffic / (per_launch * BYTES_PER_FACE) if traffic else None,
        "issued_fp64_tflops": per_launch * ISSUED_FP64_OPS_PER_FACE / avg_launch_s / 1e12,
        "hbm_algorithmic_gbs": gbs,
        "avg_launch_ms": avg_launch_s * 1e3,
        "faces_per_launch": per_launch,
        "flops_per_face": FLOPS_PER_FACE,
        "bytes_per_face": BYTES_PER_FACE,
    }


def quadrants_section(P, K, W):
    """The four-quadrant problem at full size (2049 x 2049, 7 levels, 33 x 33
    coarse, eps 1e-3, CFL 0.4): advanced to its first output time, then a
    window of K steps."""
    q, dx = P.ic_quadrants_2d(Q_RES, 0.5)
    S = P.Solver(2, Q_LEVELS, Q_NXC, 1.4, dx, ny_coarse=Q_NXC, x0=0.0, bc_kind=P.BC_TRANSMISSIVE)
    pts = P.print_times(Q_T_FINAL, Q_NUM_PRINT)
    S.set_schedule(pts)
    S.upload_q(q)
    t0 = time.perf_counter()
    adv = 0
    while S.t < pts[1]:
        n, _ = S.run(Q_TOL, Q_CFL, pts[1], 256)
        adv += n
        if n == 0:
            break
    adv_s = time.perf_counter() - t0
    S.run(Q_TOL, Q_CFL, Q_T_FINAL, W)
    t_start = S.t
    steps, upd, ms = S.run_timed(Q_TOL, Q_CFL, Q_T_FINAL, K)
    return {"workload": "2-D four-quadrant Riemann problem AWCM, 2049x2049 finest (7 levels, 33x33 coarse), eps=1e-3, "
                        "WENO-Z5+Roe, SSP-RK3 CFL 0.4, FP64",
            "advance_steps": adv, "advance_wall_s": adv_s, "t_start": t_start, "t_end": S.t, "steps": steps,
            "updates": upd, "ms_per_step": ms / max(steps, 1), "updates_per_s": upd / (ms / 1e3),
            "mean_active_points": upd / max(steps, 1)}


def advect_section(P, iters=100, cpu_iters=5, cpu=True):
    """The scalar box-advection driver (Wavelet/main_2d_advect.cpp) in its own
    configuration -- 449 x 449 (6 levels, 15 x 15 coarse), eps 1e-4, dt = 0.4
    dx -- for its 100 iterations in one awcm_run_scalar call (host wall clock
    around the synchronous call, after a warm-up on another handle), with the
    oracle port's host time per iteration from the same start beside it."""
    L, nxc, eps = 6, 15, 1e-4  # main_2d_advect.cpp:10-15
    res = (nxc - 1) * 2 ** (L - 1) + 1
    q, dx = P.ic_box_advection_2d(res)
    dt = 0.4 * dx

    def fresh():
        S = P.Solver(2, L, nxc, 1.4, dx, ny_coarse=nxc, x0=0.0, bc_kind=P.BC_TRANSMISSIVE)
        S.upload_scalar(q)
        return S

    fresh().run_scalar(eps, dt, 2)
    S = fresh()
    t0 = time.perf_counter()
    S.run_scalar(eps, dt, iters)
    wall = time.perf_counter() - t0
    m, _ = S.masks()
    out = {"workload": "main_2d_advect.cpp's loop: box advection 449x449 (6 levels, 15x15 coarse), eps 1e-4, "
                       "dt = 0.4 dx, WENO5 right reconstructions + AWCM central differences, SSP-RK3, FP64",
           "iterations": iters, "wall_s": wall, "ms_per_iteration": 1e3 * wall / iters,
           "active_points_end": int(m[1, 3:-3, 3:-3].sum())}
    if cpu:
        # the CPU port (oracle) from the same start: a reported baseline only
        import oracle_lib as O
        xl, xu, _ = O.box_grid(res)
        R = O.Solver2D("orc", L, nxc, nxc, 1.4, 0, xl, xu, xl, xu)
        R.set_scalar(q)
        t0 = time.perf_counter()
        R.advect(eps, dt, cpu_iters)
        c = (time.perf_counter() - t0) / cpu_iters
        out.update(cpu_port_ms_per_iteration=1e3 * c, cpu_port_iterations=cpu_iters,
                   cpu_port_threads=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                   gpu_over_cpu_port=c / (wall / iters))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--only", type=float, default=None,
                    help="advance untimed to this time and time one window there (for profiling)")
    ap.add_argument("--no-full", action="store_true", help="advance untimed to the windows instead of the timed run")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of each CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 1-D and adjoint sections")
    ap.add_argument("--no-c2", action="store_true", help="skip the four-quadrant window")
    ap.add_argument("--batch-steps", type=int, default=40,
                    help="steps per instance of the batched-vs-sequential comparison")
    ap.add_argument("--cpu-sample", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--c4-cpu-sample", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--c4-cpu-workers", type=int, default=1, help=argparse.SUPPRESS)
    ap.add_argument("--hbm-copy", action="store_true")
    ap.add_argument("--save-snapshot", default=None,
                    help="with --only T: advance to T, save the device state to this .npz and exit")
    ap.add_argument("--load-snapshot", default=None,
                    help="start the one window from a state saved by --save-snapshot (for short profiled runs)")
    args = ap.parse_args()
    if args.cpu_sample:
        print(json.dumps(cpu_sample(args.cpu_sample, args.cpu_budget)))
        return
    if args.c4_cpu_sample:
        print(json.dumps(c4_cpu_sample(args.c4_cpu_sample, args.c4_cpu_workers)))
        return
    if args.hbm_copy:
        hbm_copy()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build runs every solve on a single MI355X); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    P = ensure_built()
    S = dmr_solver(P)
    if args.load_snapshot:
        z = np.load(args.load_snapshot)
        args.only = float(z["t"])
    marks = (args.only,) if args.only is not None else WINDOWS
    headline = args.only if args.only is not None else HEADLINE
    full = None
    if args.load_snapshot:
        snaps = {args.only: (z["q"], z["m0"], float(z["t"]), int(z["cnt"]))}
    elif args.only is None and not args.no_full:
        full, snaps = full_run(S, T_FINAL, marks)
    else:
        snaps = {}
        for w in marks:
            while S.t < w:
                n, _ = S.run(TOL, CFL, w, 256)
                if n == 0:
                    break
            snaps[w] = snapshot(S)
    if args.save_snapshot:
        q, m0, t, cnt = snaps[headline]
        np.savez(args.save_snapshot, q=q, m0=m0, t=t, cnt=cnt)
        print(json.dumps({"saved": args.save_snapshot, "t": t, "cnt": cnt}))
        return
    wins, starts = {}, {}
    for w in marks:
        wins[w], starts[w] = window(P, S, snaps[w], args.steps, args.warmup)
    head = wins[headline]
    steps, ms = head["steps"], head["device_ms"]
    value = head["updates_per_s"]

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms / max(steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (double Mach reflection initial condition, advanced on the GPU)",
        "config": {"workload": "2-D double Mach reflection AWCM, 4097x1025 finest (7 levels, 65x17 coarse), "
                               "eps=1e-3, WENO-Z5+Roe, SSP-RK3 CFL 0.4, FP64; value = the window of --steps steps "
                               f"after --warmup steps from t~{headline}",
                   "t_start": head["t_start"], "t_end": head["t_end"], "updates": head["updates"],
                   "mean_active_points": head["mean_active_points"]},
        "roofline": faces_roofline(head),
        "step_hbm": {"algorithmic_bytes_per_step": head["algorithmic_bytes_per_step"],
                     "achieved_gbs": head["algorithmic_gbs"], "peak_gbs": HBM_PEAK_GBS, "frac": head["hbm_frac"]},
        "breakdown_ms_per_step": head["breakdown_ms_per_step"],
        "windows": {f"t~{w}": wins[w] for w in marks},
    }
    if full:
        out["full_run"] = dict(full, workload="the whole DMR run, t = 0 .. 0.2, HIP-event device time over "
                                              "chunks of 32 graph-replayed steps")
    if not args.no_c2 and args.only is None:
        out["quadrants"] = quadrants_section(P, args.steps, args.warmup)
    if not args.no_batch and args.only is None:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["batched"] = batched_section(P, args.batch_steps, 0 if args.no_cpu else threads)
        out["adjoint"] = adjoint_section(P)
        out["scalar_advect"] = advect_section(P, cpu=not args.no_cpu)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cb = cpu_baseline(starts[headline], args.cpu_budget, threads)
        out["cpu_baseline"] = cb
        if "value" in cb and cb["value"] > 0:
            out["gpu_over_cpu"] = value / cb["value"]
        c1 = cpu_baseline(starts[headline], args.cpu_budget, 1)
        out["cpu_baseline_1core"] = c1
        if "value" in c1 and c1["value"] > 0:
            out["gpu_over_cpu_1core"] = value / c1["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
