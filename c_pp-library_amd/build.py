"""Build the HIP library libawcm_hip.so for gfx950 (in place, next to this file).

    python c_pp-library_amd/build.py

Flags that matter for correctness:
  -ffp-contract=off   no FMA contraction: the FP64 operation order is the
                      reference's, which bit-exact parity needs (hipcc would
                      otherwise contract a*b+c into fma).
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed-FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = ["csrc/awcm1d.hip", "csrc/adjoint1d.hip", "csrc/runtime1d.hip", "csrc/awcm2d.hip", "csrc/capi.hip", "csrc/noalloc1d.hip", "csrc/problems.cpp"]
OUT = os.path.join(HERE, "libawcm_hip.so")
ARCH = os.environ.get("AWCM_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose=False):
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-fno-fast-math", "-Wno-unused-value", "-Wno-unused-result",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "include"),
           "-I", os.path.join(HERE, "csrc"),
           "-o", OUT + ".tmp"] + [os.path.join(HERE, s) for s in SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    # the host pass does not know the packed-FP32 feature flag; that note is harmless
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if err.strip():
        print(err, file=sys.stderr)
    if r.returncode != 0:
        raise subprocess.CalledProcessError(r.returncode, cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
    print("built", OUT)
