"""Compact digests of a solver state (test infrastructure): SHA-256 of the bits
of the AoS ghosted field, the five masks (uint8) and the four level fields (as
int8), plus the number of mask0 points.  tests/golden/make_digests.py writes
them from the reference build; the GPU tests compute the same from the device
state."""
import hashlib

import numpy as np


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:32]


def digest(q, masks, levels):
    m = np.asarray(masks, dtype=np.uint8)
    return {"q": sha(np.asarray(q, dtype=np.float64)), "masks": sha(m),
            "levels": sha(np.asarray(levels).astype(np.int8)), "n_mask0": int(m[1].sum())}


def same(d, want):
    """Fields of digest d that differ from want (q, masks, levels, n_mask0)."""
    return [k for k in ("q", "masks", "levels", "n_mask0") if d[k] != want[k]]
