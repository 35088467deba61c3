"""Seeded synthetic workloads.

The batched shock-tube workload: n independent 1-D Riemann problems on [0, 1]
with gamma = 1.4, drawn from an in-repo splitmix64 generator (a double is the
top 53 bits of a draw times 2**-53) so the instances are identical on every
platform:

    rho_L, rho_R ~ U[0.1, 1];  p_L, p_R ~ U[0.05, 1];  u ~ U[-0.5, 0.5]
    (one velocity for both states);  interface x0 ~ U[0.3, 0.7];
    J_max ~ U{9..12}, i.e. 6..9 levels over a 17-point coarsest grid
    (J_min = 4; 513..4097 finest points);  eps in {1e-3, 1e-4};
    CFL 0.5 to t = 0.15.

Draw order per instance: rho_L, rho_R, p_L, p_R, u, x0, J_max, eps.
"""
import numpy as np

import pyawcm

MASK64 = (1 << 64) - 1
C4_SEED = 0x5EED0C4
C4_T_FINAL = 0.15
C4_CFL = 0.5
C4_NUM_PRINT = 41  # output times of the 1-D driver (main_1d_euler.cpp:25)


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * 2.0 ** -53)


def shock_tube_instances(n=512, seed=C4_SEED):
    """[(num_levels, rho_L, rho_R, p_L, p_R, u, x0, eps)] for n instances."""
    g = SplitMix64(seed)
    out = []
    for _ in range(n):
        rl, rr = g.uniform(0.1, 1.0), g.uniform(0.1, 1.0)
        pl, pr = g.uniform(0.05, 1.0), g.uniform(0.05, 1.0)
        u = g.uniform(-0.5, 0.5)
        x0 = g.uniform(0.3, 0.7)
        jmax = 9 + min(int(g.uniform() * 4), 3)
        eps = 1e-3 if g.uniform() < 0.5 else 1e-4
        out.append((jmax - 3, rl, rr, pl, pr, u, x0, eps))
    return out


def shock_tube_batch(n=512, seed=C4_SEED, t_final=C4_T_FINAL):
    """A pyawcm.Batch1D holding the instances with their initial states, plus
    the instance list and the initial states (for checks)."""
    inst = shock_tube_instances(n, seed)
    params, states = [], []
    for (L, rl, rr, pl, pr, u, x0, eps) in inst:
        nx = 16 * 2 ** (L - 1) + 1
        q, dx, _ = pyawcm.ic_riemann_1d(nx, 0.0, 1.0, x0, (rl, u, pl), (rr, u, pr))
        params.append((L, 17, 1.4, dx, eps, C4_CFL, t_final))
        states.append(q)
    b = pyawcm.Batch1D(params, pyawcm.print_times(t_final, C4_NUM_PRINT))
    for k, q in enumerate(states):
        b.upload_q(k, q)
    return b, inst, params, states
