"""The reference-named C++ class API (include/awcm_api.h through include/compat/),
host side.

tests/native/api_probe.cpp uses only class names, members and methods that the
reference defines too.  Built against the reference's own headers it printed
tests/golden/api_probe_host.txt (tests/golden/make_api_probe_golden.sh); built
against this library it must print the same text.  That covers bit-for-bit:
  VECTOR_3D/4D arithmetic, IDEAL_GAS_1D/2D, EULER_EQUATION_1D/2D,
  ROE_SOLVER_1D/2D, WENOZ5 and WENO5 reconstructions, the WENO-Z and WENO5-JS
  Euler face fluxes (1-D, 2-D x/y) over random, smooth and uniform stencils,
  MRA_GRID_INT_1D/2D level ranges, the four problems' Initialize (grids and
  fields with ghosts), the three boundary managers, AWCM_1D/2D
  differentiation stencils, constructor state and PostProcessEuler, the files
  Print() writes (names and bytes), and OMP_MANAGER's reductions.
The GPU half (PrepareEuler and the drivers' time loop) is tests/test_gpu_api.py.

Where the reference tree is present, the probe is also rebuilt from it and run
live, and the shipped drivers (Wavelet/main_1d_euler.cpp, main_2d_euler.cpp,
main_2d_euler(dmach).cpp, main_2d_euler(quad).cpp and the scalar
main_2d_advect.cpp) are compiled and linked against this library with only
`void main` -> `int main` and sprintf_s -> snprintf changed.
(Wavelet/main_1d_advect.cpp is not among them: it calls a six-argument
AWCM_1D::DifferentiateAtLeftBoundary that the reference does not define, so
it does not build against the reference either.)
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
PROBE = os.path.join(NATIVE, "_build", "awcm_api_probe")
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF = "/root/reference"
REF_PROBE = os.path.join(ROOT, "oracle", "_ref", "awcm_api_probe_ref")
REF_PROBE_ADVECT = os.path.join(ROOT, "oracle", "_ref", "awcm_api_probe_advect_ref")
DRIVERS = ["main_1d_euler", "main_2d_euler", "main_2d_euler(dmach)", "main_2d_euler(quad)", "main_2d_advect"]


def have_ref():
    return os.path.exists(os.path.join(REF, "WENOZ.h"))


@pytest.fixture(scope="module")
def probe():
    import pyawcm
    if not os.path.exists(pyawcm.LIB_PATH):
        pytest.skip("HIP library not built (python c_pp-library_amd/build.py)")
    subprocess.run(["make", "-s", "-C", NATIVE], check=True)
    return PROBE


def golden(name):
    with open(os.path.join(GOLDEN, f"api_probe_{name}.txt")) as f:
        return f.read().splitlines()


def diff_lines(got, want):
    bad = [(w, g) for w, g in zip(want, got) if w != g]
    if len(got) != len(want):
        bad.append((f"{len(want)} lines", f"{len(got)} lines"))
    return bad


def run_host(exe, tmp_path):
    out = subprocess.run([exe, "host", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return out.stdout.splitlines()


def test_host_classes_match_reference_output(probe, tmp_path):
    got = run_host(probe, tmp_path)
    bad = diff_lines(got, golden("host"))
    assert not bad, "differs from the reference build (expected, got): " + repr(bad[:8])


def test_print_files_named_as_reference(probe, tmp_path):
    run_host(probe, tmp_path)
    names = set(os.listdir(tmp_path))
    # FIELD_UNIFORM_1D::Print() numbers "<name>%4.4d", the 2-D one "<name>_%4.4d"
    assert {"nm0000", "nm0001", "Q_0000", "m1v", "m2"} <= names


@pytest.mark.skipif(not have_ref(), reason="reference tree not present")
def test_golden_is_what_the_reference_prints(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref-probe"], check=True)
    got = run_host(REF_PROBE, tmp_path)
    assert not diff_lines(got, golden("host"))
    for mode, arg in (("prepare", []), ("manager", []), ("loop", ["40"])):
        out = subprocess.run([REF_PROBE, mode] + arg, capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stderr
        assert not diff_lines(out.stdout.splitlines(), golden(mode)), mode
    out = subprocess.run([REF_PROBE_ADVECT, "40"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert not diff_lines(out.stdout.splitlines(), golden("advect")), "advect"


def patched_driver(name, dst):
    """The shipped driver with the two changes it needs to build with g++."""
    with open(os.path.join(REF, "Wavelet", name + ".cpp")) as f:
        src = f.read()
    src = re.sub(r"^void main\(\)", "int main()", src, flags=re.M)
    src = re.sub(r"sprintf_s\((\w+),", r"snprintf(\1, sizeof \1,", src)
    path = os.path.join(dst, name + ".cpp")
    with open(path, "w") as f:
        f.write(src)
    return path


@pytest.mark.skipif(not have_ref(), reason="reference tree not present")
@pytest.mark.parametrize("name", DRIVERS)
def test_shipped_driver_builds_against_library(name, tmp_path, probe):
    src = patched_driver(name, str(tmp_path))
    exe = os.path.join(str(tmp_path), "driver")
    cmd = ["g++", "-std=c++17", "-O1", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wno-unknown-pragmas",
           "-Wno-unused-variable", "-Wno-unused-but-set-variable",
           "-I", os.path.join(ROOT, "include", "compat"), "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "c_pp-library_amd", "include"), src,
           "-L", os.path.join(ROOT, "c_pp-library_amd"), "-lawcm_hip", "-o", exe]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-4000:]
    assert os.path.exists(exe)


def test_compat_headers_cover_driver_includes():
    """Every header name the three drivers and the problem files include."""
    want = ["OMP_MANAGER.h", "STOP_WATCH.h", "WENO.h", "WENOZ.h", "EULER_EQUATION.h", "ARRAY.h", "ARRAY_1D.h",
            "AWCM_2D.h", "Wavelet/AWCM_1D.h", "Problems/1D/PROBLEM_SHOCKTUBE.h",
            "Problems/1D/PROBLEM_ENTROPY_INTERACTING.h", "Problems/2D/PROBLEM_DOUBLEMACH.h",
            "Problems/2D/PROBLEM_QUADRANTS.h", "Problems/2D/PROBLEM_BOX_ADVECTION.h", "Wavelet/WAVELET_2D.h"]
    compat = os.path.join(ROOT, "include", "compat")
    missing = [w for w in want if not os.path.exists(os.path.join(compat, w))]
    assert not missing, missing
