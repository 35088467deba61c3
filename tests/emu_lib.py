"""Build and load the host emulation of the 1-D HIP kernels
(tests/native/emu1d.cpp: k1_solve with tape recording, k1_adjoint, and the
reverse sweep's grid kernels, compiled unchanged for the CPU with g++).

Test infrastructure only: the product never imports this module.
"""
import ctypes as C
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_dp = C.POINTER(C.c_double)


def build(out, threads=1, extra=()):
    """Compile the emulation into the shared library `out` and load it."""
    pkg = os.path.join(ROOT, "c_pp-library_amd")
    inc = os.path.join(ROOT, "tests", "native", "emu_mt" if threads > 1 else "emu")
    subprocess.run(["g++", "-std=c++20" if threads > 1 else "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC",
                    "-shared", "-w", f"-DEMU_THREADS={threads}", *extra,
                    "-I", inc, "-I", os.path.join(pkg, "include"), "-I", os.path.join(pkg, "csrc"),
                    "-I", os.path.join(ROOT, "include"), "-x", "c++",
                    os.path.join(ROOT, "tests", "native", "emu1d.cpp"), "-o", out, "-lpthread"], check=True)
    E = C.CDLL(out)
    V = C.c_void_p
    E.emu_create.restype = V
    E.emu_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, C.c_int]
    E.emu_destroy.argtypes = [V]
    E.emu_set_q.argtypes = [V, _dp]
    E.emu_get_q.argtypes = [V, _dp]
    E.emu_record.argtypes = [V, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
    E.emu_record.restype = C.c_int
    E.emu_record_from.argtypes = [V, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int]
    E.emu_record_from.restype = C.c_int
    E.emu_replay.argtypes = [V]
    E.emu_adjoint.argtypes = [V, _dp, C.c_int, _dp]
    E.emu_status.argtypes = [V]
    E.emu_set_masks_in_hbm.argtypes = [C.c_int]
    E.emu_set_lds_auto.argtypes = [C.c_int]
    E.emu_set_poison.argtypes = [C.c_int]
    E.emu_check_weights.restype = C.c_int
    return E
