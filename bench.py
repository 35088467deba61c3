#!/usr/bin/env python3
"""Headline benchmark: 2-D double Mach reflection AWCM forward pass on one MI355X.

Workload (the configuration the headline metric is quoted on): Mach-10 shock
at 60 degrees on [0,4]x[0,1], finest grid 4097 x 1025 (7 levels, 65 x 17
coarse), pre-shock state (1.4, 0, 0, 1), eps = 1e-3, WENO-Z5 + Roe, SSP-RK3 at
CFL 0.4, output schedule of 21 times to t = 0.2, FP64.

A step is one full AWCM time step (PrepareEuler + CFL dt + three SSP-RK3
stages + post-processing) over the device-resident state.  The solution is
first advanced to --advance (untimed) so the measured steps see a developed
flow, then --warmup steps run untimed and --steps steps are timed with HIP
events on the solver's stream (no host synchronisation inside the window).

value = active-point RK3 updates per second = sum over the timed steps of
|mask0| / device time.  The CPU baseline resumes the reference build (or the
oracle port) from the exact state the timed window started from and times a
bounded number of the same steps on the host cores.

Prints one JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
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

# Algorithmic cost of the dominant kernel, k2_stage_faces (one WENO-Z/Roe face
# flux per thread, each face of the stage once).  FP64 operations of one face
# flux as the reference writes it (+ - * / sqrt, WENOZ.h:236-320 with
# ROESOLVER.h / EULER_EQUATION.h), counted mechanically: 1202.
FLOPS_PER_FACE = 1202
# Bytes per face: the face's own state (32 B, neighbours' states are the
# neighbouring faces' own reads), its 4 B list entry, and the 32 B flux written.
BYTES_PER_FACE = 32 + 4 + 32
# Measured on this workload at t = 0.02 with rocprofv3 (profiles/
# k2_stage_faces_dmr_pmc.md): HBM-side bytes per k2_stage_faces launch =
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors from a 1 GiB device copy), and
# issued FP64 lane operations per face (ADD + MUL + TRANS + 2 FMA, x 64 lanes).
TRAFFIC_BYTES_PER_LAUNCH = 94705466
ISSUED_FP64_OPS_PER_FACE = 2143


def ensure_built():
    import pyawcm
    if not os.path.exists(pyawcm.LIB_PATH):
        subprocess.run([sys.executable, os.path.join(ROOT, "c_pp-library_amd", "build.py"), "-q"], check=True)
    return pyawcm


def cpu_baseline(q, m0, t, cnt, budget_s):
    """Time the same steps on the host from the GPU state (reference build if present)."""
    try:
        import oracle_lib as O
    except Exception as e:  # pragma: no cover
        return {"error": f"oracle unavailable: {e}"}
    kind, prefix = ("reference", "ref") if O.have_ref() else ("port", "orc")
    try:
        s = O.Solver2D(prefix, LEVELS, NXC, NYC, 1.4, 1, 0.0, 4.0, 0.0, 1.0)
    except FileNotFoundError as e:
        return {"error": str(e)}
    s.set_q(q)
    s.set_mask0(m0)
    s.t[0] = t
    s.cnt.value = cnt
    pts = O.print_times(T_FINAL, NUM_PRINT)
    upd, elapsed, steps = 0, 0.0, 0
    while elapsed < budget_s and steps < 50:
        t0 = time.perf_counter()
        s.step(TOL, CFL, pts)
        elapsed += time.perf_counter() - t0
        steps += 1
        m, _ = s.masks()
        upd += int(m[1, 3:-3, 3:-3].sum())
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": upd / elapsed, "unit": "updates/s", "cores": threads, "kind": kind,
            "sample": f"{steps} DMR steps resumed from the GPU state at t={t:.6f} "
                      f"({elapsed:.1f} s wall, OpenMP {threads} threads, {prefix} build)"}


def batched_section(P, cmp_steps):
    """512 seeded 1-D shock tubes (uneven 513..4097-point grids, eps 1e-3/1e-4,
    CFL 0.5, t = 0.15) in one graph-replayed persistent launch with the
    autotuned shape, and the same work as one plain launch per instance."""
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
    return {
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--advance", type=float, default=0.02, help="untimed advance to this time first")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 1-D and adjoint-forward sections")
    ap.add_argument("--batch-steps", type=int, default=40,
                    help="steps per instance of the batched-vs-sequential comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build runs every solve on a single MI355X); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    P = ensure_built()
    q0, dx = P.ic_double_mach_2d(RES, RHO_PRE)
    S = P.Solver(2, LEVELS, NXC, 1.4, dx, ny_coarse=NYC, x0=0.0, bc_kind=P.BC_DOUBLE_MACH)
    S.set_schedule(P.print_times(T_FINAL, NUM_PRINT))
    S.upload_q(q0)
    t_adv0 = time.perf_counter()
    adv_steps = 0
    while S.t < args.advance:
        n, _ = S.run(TOL, CFL, args.advance, 256)
        adv_steps += n
        if n == 0:
            break
    adv_s = time.perf_counter() - t_adv0
    S.run(TOL, CFL, T_FINAL, args.warmup)
    S.synchronize()
    # state at the start of the timed window: replayed for the per-kernel timing
    # pass and handed to the CPU baseline
    t_start, cnt_start = S.t, S.cnt
    q_start = S.download_q()
    m_start = S.masks()[0][1].copy()

    # timed window: whole-step HIP graph replays, no host synchronisation
    steps, updates, ms = S.run_timed(TOL, CFL, T_FINAL, args.steps)
    t_end = S.t

    # per-kernel HIP-event timing over the same window (state restored bit-exactly)
    S.upload_q(q_start)
    S.set_mask0(m_start)
    S.t, S.cnt = t_start, cnt_start
    S.profile(True)
    steps_p, updates_p, ms_p = S.run_timed(TOL, CFL, T_FINAL, args.steps)
    w_ms, w_launch, w_pts = S.profile_read(P.KERNEL_STAGE_WENO)
    c_ms, c_launch, c_pts = S.profile_read(P.KERNEL_STAGE_COARSE)
    p_ms, p_launch, _ = S.profile_read(P.KERNEL_PREPARE)
    S.profile(False)
    assert steps_p == steps and updates_p == updates and S.t == t_end, "profiled replay diverged"

    value = updates / (ms / 1e3) if ms > 0 else 0.0
    # dominant kernel: WENO-Z/Roe face fluxes of a stage
    avg_launch_s = (w_ms / w_launch) / 1e3 if w_launch else float("nan")
    faces_per_launch = w_pts / w_launch if w_launch else 0.0
    achieved_tflops = faces_per_launch * FLOPS_PER_FACE / avg_launch_s / 1e12 if w_launch else 0.0
    achieved_gbs = faces_per_launch * BYTES_PER_FACE / avg_launch_s / 1e9 if w_launch else 0.0

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
                               "eps=1e-3, WENO-Z5+Roe, SSP-RK3 CFL 0.4, FP64",
                   "t_start": t_start, "t_end": t_end, "eager_ms_per_step": ms_p / max(steps, 1), "advance_steps": adv_steps,
                   "advance_wall_s": round(adv_s, 3), "updates": updates,
                   "mean_active_points": updates / max(steps, 1)},
        "roofline": {
            "bound": "mfma",
            "compute": "FP64 vector ALU: the kernel issues no MFMA; MI355X's FP64 vector and matrix peaks are both "
                       "78.6 TFLOP/s, and bit-exact parity forbids FMA contraction",
            "kernel": "k2_stage_faces (WENO-Z/Roe face fluxes, each face of a stage once)",
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "traffic": TRAFFIC_BYTES_PER_LAUNCH,
            "traffic_over_algorithmic": TRAFFIC_BYTES_PER_LAUNCH / (faces_per_launch * BYTES_PER_FACE)
            if faces_per_launch else None,
            "issued_fp64_tflops": faces_per_launch * ISSUED_FP64_OPS_PER_FACE / avg_launch_s / 1e12 if w_launch else 0.0,
            "hbm_achieved_gbs": achieved_gbs,
            "hbm_frac": achieved_gbs / HBM_PEAK_GBS,
            "avg_launch_ms": avg_launch_s * 1e3,
            "faces_per_launch": faces_per_launch,
            "flops_per_face": FLOPS_PER_FACE,
            "bytes_per_face": BYTES_PER_FACE,
        },
        "breakdown_ms_per_step": {
            "stage_faces": w_ms / max(steps, 1),
            "stage_update": c_ms / max(steps, 1),
            "prepare_euler": p_ms / max(steps, 1),
        },
    }
    if not args.no_batch:
        out["batched"] = batched_section(P, args.batch_steps)
        out["adjoint"] = adjoint_section(P)
    if not args.no_cpu:
        cb = cpu_baseline(q_start, m_start, t_start, cnt_start, args.cpu_budget)
        out["cpu_baseline"] = cb
        if "value" in cb and cb["value"] > 0:
            out["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
