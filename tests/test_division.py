"""The device's quotients by a shared divisor (struct Div, awcm/physics.h).

Host side: tests/native/markstein_exhaustive.c emulates the device sequence
exactly in p-bit binary formats and checks every pair of significands for
p = 4..12: the reciprocal step gives RN(1/b) from either bracket of 1/b with
exact residuals, and the corrected quotient is RN(a/b).  A build that skips the
reciprocal correction must fail the same check.

The device side (test_gpu_division.py) compares the kernels' quotient with the
host's IEEE division bit for bit.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "markstein_exhaustive.c")


@pytest.fixture(scope="module")
def exes(tmp_path_factory):
    d = tmp_path_factory.mktemp("markstein")
    good, planted = str(d / "mex"), str(d / "mex_planted")
    subprocess.run(["gcc", "-O2", "-Wall", SRC, "-o", good], check=True)
    subprocess.run(["gcc", "-O2", "-Wall", "-DPLANT_NO_RECIP_CORRECTION", SRC, "-o", planted], check=True)
    return good, planted


def _run(exe, p):
    r = subprocess.run([exe, str(p)], capture_output=True, text=True)
    m = re.search(r"pairs=(\d+) recip_bad=(\d+) quot_bad=(\d+) inexact=(\d+)", r.stdout)
    assert m, r.stdout + r.stderr
    return r.returncode, [int(x) for x in m.groups()]


@pytest.mark.parametrize("p", [4, 5, 6, 7, 8, 9, 10, 11, 12])
def test_shared_divisor_quotient_is_correctly_rounded(exes, p):
    rc, (pairs, recip_bad, quot_bad, inexact) = _run(exes[0], p)
    assert pairs == 4 ** (p - 1)
    assert (recip_bad, quot_bad, inexact) == (0, 0, 0)
    assert rc == 0


def test_uncorrected_reciprocal_is_caught(exes):
    rc, (_, recip_bad, quot_bad, _) = _run(exes[1], 10)
    assert rc != 0 and recip_bad > 0 and quot_bad > 0
