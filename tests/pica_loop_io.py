"""Input / output files of tests/cxx/pica_loop.cpp (a pica caller's loop) and the oracle run it is
checked against."""
import numpy as np

import oracle as O
from helpers import courant_dt, plasma, random_particles


def write_input(path, ncell, ncps, x, u, w):
    with open(path, "wb") as f:
        np.array(ncell, np.int32).tofile(f)
        np.array(ncps, np.int32).tofile(f)
        np.array([x.shape[0]], np.int64).tofile(f)
        np.concatenate([x, u, w[:, None]], 1).astype(np.float64).tofile(f)


def read_output(path, ncell):
    raw = np.fromfile(path, np.uint8)
    n = int(raw[:8].view(np.int64)[0])
    p = raw[8:8 + n * 56].view(np.float64).reshape(n, 7)
    f = raw[8 + n * 56:].view(np.float64).reshape((9,) + tuple(c + 4 for c in ncell))
    return p, f


def case(seed=21, ncell=(8, 8, 12), ncps=(4, 4, 4), ppc=4):
    """particles with unique weights (to match them as a set) and the same run on the oracle"""
    rng = np.random.default_rng(seed)
    dt = courant_dt()
    _, w0 = plasma(dt, ppc)
    x, u = random_particles(rng, ncell, ppc, uth=0.05, drift=0.1)
    w = w0 * (1.0 + np.arange(x.shape[0]) * 2.0 ** -20)
    return ncell, ncps, x, u, w, dt


def oracle_run(ncell, ncps, x, u, w, dt, nsteps):
    og = O.Grid.periodic(ncell)
    oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    for _ in range(nsteps):
        O.step(og, oe, dt, O.CIC, O.DIRECT)
    return og, oe
