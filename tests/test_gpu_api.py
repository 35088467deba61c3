"""The reference-named C++ class API on the GPU: AWCM_1D/2D::PrepareEuler runs
through awcm_prepare_euler, and the drivers' time loop runs through the class
API (host stages, device PrepareEuler every step).

tests/native/_build/awcm_api_probe is the same source that, built against the
reference's headers and run on the CPU, printed tests/golden/api_probe_*.txt.
Bar: identical text, i.e. every field (ghosts included), mask and level field
bit-identical, and t and dt identical at every step:
  prepare  two consecutive PrepareEuler calls on Sod and Shu-Osher (1-D, 6
           levels, 481 points), on Sod as main_1d_euler.cpp configures it (11
           levels, 15361 points: masks beyond one workgroup's LDS, so the
           device keeps them in HBM) and on the quadrants and double-Mach
           problems (2-D, 4 levels, 65x65 and 129x33), each 2-D one followed
           by WAVELET_2D::InterpolateAllGridEuler from mask1, then from mask0
  loop     40 steps of main_1d_euler.cpp's loop (WENO-Z face-flux pair on Sod
           and Shu-Osher; WENO5-JS ComputedF, as the shipped driver, on Sod at
           481 points and in the driver's own 15361-point configuration) and of
           main_2d_euler(quad|dmach).cpp's loop, each 2-D loop ending with
           the whole grid from mask0 as main_2d_euler.cpp:334 does
  advect   the scalar path: 40 steps of main_2d_advect.cpp's loop (box
           advection, AWCM_2D::Prepare / PostProcess, WENO5 right
           reconstructions) on 65x65 and 10 in the driver's own 449x449
           configuration, each ending with WAVELET_2D::Interpolate from mask0
           onto every point (main_2d_advect.cpp:219) and from mask1 onto mask3
           (tests/native/_build/awcm_api_probe_advect)
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "tests", "native", "_build", "awcm_api_probe")
PROBE_ADVECT = os.path.join(ROOT, "tests", "native", "_build", "awcm_api_probe_advect")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def run(mode, *args, exe=PROBE):
    import pyawcm
    pyawcm.lib()  # raises if the HIP library is not built
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    print(f"api probe: {os.path.basename(exe)} {mode} {' '.join(args)}", flush=True)
    out = subprocess.run([exe, mode, *args], capture_output=True, text=True, timeout=110)
    assert out.returncode == 0, f"probe {mode} failed ({out.returncode}): {out.stderr[-2000:]}"
    return out.stdout.splitlines()


def compare(got, name):
    with open(os.path.join(GOLDEN, f"api_probe_{name}.txt")) as f:
        want = f.read().splitlines()
    bad = [(w, g) for w, g in zip(want, got) if w != g]
    assert len(got) == len(want), f"{len(got)} lines, expected {len(want)}; first differences {bad[:4]}"
    assert not bad, f"differs from the reference (expected, got): {bad[:6]}"


def test_prepare_euler_matches_reference():
    compare(run("prepare"), "prepare")


def test_manager_statics_match_reference():
    """AWCM_MANAGER_1D/2D's static mask algebra through the class API."""
    compare(run("manager"), "manager")


def test_driver_loop_matches_reference():
    compare(run("loop", "40"), "loop")


def test_advect_loop_matches_reference():
    compare(run("40", exe=PROBE_ADVECT), "advect")
