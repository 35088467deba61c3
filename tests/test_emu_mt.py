"""CPU check of the 1-D kernels' barrier discipline with a workgroup of real threads.

tests/native/emu1d.cpp (the device code of k1_solve and k1_adjoint, compiled
unchanged) is built against tests/native/emu_mt/hip/hip_runtime.h: every GPU
thread of the one workgroup is a host thread, __syncthreads a barrier, LDS
shared memory, __shfl_xor an exchange inside a 64-thread wave.  Built with
ThreadSanitizer and driven by tests/native/emu1d_mt_main.cpp (a recorded Sod
run on the adaptive grid, its frozen replay, the reverse sweep with two
recompute segment lengths), any two threads touching the same memory between
barriers without atomics is reported.

One pattern is suppressed: set1, the mask update of the reconstruction check,
adjacent zone and points for differentiation, where several threads store the
same byte value 1 into one mask entry in the same pass (all of the
reference's mask updates only ever add points).  Every other access must be
race-free, and the run must also reproduce its recording and give the same
gradient for both segment lengths.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exes(tmp_path_factory):
    d = tmp_path_factory.mktemp("emu_mt")
    pkg = os.path.join(ROOT, "c_pp-library_amd")
    out = {}
    for t in (64, 256):
        exe = str(d / f"emu1d_mt_{t}")
        subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-ffp-contract=off", "-fsanitize=thread",
                        "-Wno-unknown-pragmas", f"-DEMU_THREADS={t}",
                        "-I", os.path.join(ROOT, "tests", "native", "emu_mt"), "-I", os.path.join(pkg, "include"),
                        "-I", os.path.join(pkg, "csrc"), "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "native", "emu1d.cpp"),
                        os.path.join(ROOT, "tests", "native", "emu1d_mt_main.cpp"), "-o", exe, "-lpthread"],
                       check=True)
        out[t] = exe
    supp = d / "tsan.supp"
    supp.write_text("race:set1\n")
    return out, str(supp)


@pytest.mark.parametrize("threads,levels,flux", [(64, 5, 0), (256, 4, 0), (64, 4, 1)])
def test_1d_kernels_race_free(exes, threads, levels, flux):
    """flux 1: the recording alone with the WENO5-JS stage fluxes (the
    k1_solve instantiation of the shipped driver's flux)."""
    paths, supp = exes
    env = dict(os.environ, TSAN_OPTIONS=f"suppressions={supp} halt_on_error=0")
    r = subprocess.run([paths[threads], str(levels), "8", str(flux)], capture_output=True, text=True, timeout=600,
                       env=env)
    log = r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in log, log[-4000:]
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), log[-2000:]
