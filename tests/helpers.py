"""Shared helpers of the test suite: workload set-up and error measures."""
import numpy as np

import oracle as O

C = O.C


def rel_err(a, b):
    """max |a - b| / max |b| (the per-quantity measure the parity bar is stated in)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.abs(b).max()
    if scale == 0.0:
        return float(np.abs(a).max())
    return float(np.abs(a - b).max() / scale)


def plasma(dt, ppc, omega_p_dt=0.1, step=1.0):
    """n0 with omega_p dt = omega_p_dt and the macro weight n0 h^3 / ppc (pica's CGS constants)."""
    wp = omega_p_dt / dt
    n0 = wp * wp * O.E_MASS / (4.0 * O.PI * O.E_CHARGE * O.E_CHARGE)
    return n0, n0 * step ** 3 / ppc


def courant_dt(step=1.0, courant=0.5):
    return courant * step / C


def random_particles(rng, ncell, ppc, uth=0.01, drift=0.0, step=(1.0, 1.0, 1.0)):
    """ppc particles per cell, uniform in the cell, Maxwellian u (p/(mc)), +-drift in x."""
    n0 = int(np.prod(ncell)) * ppc
    cells = np.repeat(np.arange(int(np.prod(ncell))), ppc)
    ix = cells // (ncell[1] * ncell[2])
    iy = (cells // ncell[2]) % ncell[1]
    iz = cells % ncell[2]
    x = np.stack([(ix + rng.random(n0)) * step[0], (iy + rng.random(n0)) * step[1], (iz + rng.random(n0)) * step[2]], 1)
    u = rng.standard_normal((n0, 3)) * uth
    u[:, 0] += np.where(np.arange(n0) % 2 == 0, drift, -drift)
    return x, u


def random_fields(rng, grid, e_scale, b_scale):
    """random E/B on a periodic oracle grid, ghosts filled"""
    for c in range(3):
        grid.f[c][...] = rng.standard_normal(grid.s) * e_scale
        grid.f[3 + c][...] = rng.standard_normal(grid.s) * b_scale
    for c in range(6):
        O.ghost_fill(grid.f[c])


def uneven_plasma_specs(n=64, seed=3, edges=(32, 48, 64, 96, 128), ppcs=(4, 8, 16, 32)):
    """independent plasmas of the batching form: per instance a cubic grid edge and a ppc drawn
    uniformly from the given sets, a drift speed uniform in [0, 0.2]c (returned as u = v / sqrt(1 -
    v^2)), CIC or TSC with equal odds, and its own seed"""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        v = 0.2 * rng.random()
        out.append({"edge": int(rng.choice(edges)), "ppc": int(rng.choice(ppcs)), "drift": v / np.sqrt(1.0 - v * v),
                    "shape": int(rng.integers(0, 2)), "seed": 1000 + i})
    return out
