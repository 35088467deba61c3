"""The scalar box-advection driver (Wavelet/main_2d_advect.cpp) as the tests
drive it: its two probe configurations and the per-step hash lines that
tests/native/api_probe_advect.cpp prints, from any state (oracle or device).

tests/golden/api_probe_advect.txt holds those lines as the reference build
printed them: FNV-1a over the ghosted field, the five ghosted masks
(mask_temp, mask0..3) and level0..2 (int32, no ghosts)."""
import os

import numpy as np

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "api_probe_advect.txt")

# name: (levels, coarse points per side, eps, steps in the golden text)
# "driver" is main_2d_advect.cpp's own configuration (:10-15): 449 x 449 points
CONFIGS = {"small": (4, 9, 1e-4, 40), "driver": (6, 15, 1e-4, 10)}


def golden():
    out = {}
    with open(GOLDEN) as f:
        for ln in f:
            parts = ln.split()
            if len(parts) >= 5 and parts[1].startswith("active="):
                out[parts[0]] = dict(p.split("=") for p in parts[1:])
    return out


def report(q, masks, levels):
    """The probe's line fields from a scalar field ((ny+6)*(nx+6) doubles),
    masks (5, ny+6, nx+6) and levels (>= 3, ny, nx)."""
    m = np.ascontiguousarray(masks, dtype=np.uint8)
    return {"active": str(int(m[1, 3:-3, 3:-3].sum())),
            "Q": f"{O.fnv1a(np.ascontiguousarray(q, dtype=np.float64)):016x}",
            "masks": f"{O.fnv1a(m):016x}",
            "levels": f"{O.fnv1a(np.ascontiguousarray(levels[:3], dtype=np.int32)):016x}"}


def setup(name):
    """(levels, nxc, eps, steps, res, initial field, dx, dt) of a configuration:
    BOX_ADVECTION on the finest grid, dt = 0.4 dx (main_2d_advect.cpp:36)."""
    L, nxc, eps, steps = CONFIGS[name]
    res = (nxc - 1) * 2 ** (L - 1) + 1
    q, dx = O.ic_box(res)
    return L, nxc, eps, steps, res, q, dx, 0.4 * dx
