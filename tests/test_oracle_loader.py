"""The oracle's twin of the library's seeded thermal loader (or_load_thermal), on the CPU.

Its ln and sincos are written with IEEE +, -, *, / only so that the CPU and the GPU produce the
same bits; here they are pinned against libm (math.log / sin / cos) and closed forms, and the
loader's output against what it promises: one particle per (cell, k) with the documented slot
order, positions inside their cell, Maxwellian momenta of the requested width with the two-stream
drift on even / odd particle ids.
"""
import math

import numpy as np
import pytest

import oracle as O


def test_det_log_matches_libm():
    rng = np.random.default_rng(41)
    xs = list(rng.random(2000)) + [1.0, 0.5, 0.25, 2.0 ** -53, 1.0 - 2.0 ** -53, math.sqrt(0.5),
                                   np.nextafter(math.sqrt(0.5), 1.0), 0.70710678118654746]
    for x in xs:
        if x <= 0.0:
            continue
        got, ref = O.det_log(x), math.log(x)
        assert abs(got - ref) <= 4e-16 * max(1.0, abs(ref)), (x, got, ref)
    assert O.det_log(1.0) == 0.0
    assert O.det_log(0.5) == pytest.approx(-math.log(2.0), abs=1e-16)
    assert O.det_log(2.0 ** -53) == pytest.approx(-53 * math.log(2.0), rel=1e-16)


def test_det_sincos_matches_libm():
    rng = np.random.default_rng(42)
    us = list(rng.random(2000)) + [k / 8.0 for k in range(8)] + [np.nextafter(k / 8.0, 0.0) for k in range(1, 9)]
    for U in us:
        s, c = O.det_sincos_2pi(U)
        a = 2.0 * math.pi * U
        assert abs(s - math.sin(a)) < 2e-15 and abs(c - math.cos(a)) < 2e-15, (U, s, c)
        assert abs(s * s + c * c - 1.0) < 4e-16
    # exact octant boundaries
    assert O.det_sincos_2pi(0.0) == (0.0, 1.0)
    s, c = O.det_sincos_2pi(0.25)
    assert abs(s - 1.0) < 1e-16 and abs(c) < 1e-16


@pytest.mark.parametrize("paired,ppc", [(False, 8), (True, 8), (True, 5)])
def test_loader_order_and_positions(paired, ppc):
    ncell, ncps, step = (8, 4, 12), (4, 2, 4), (1.0, 0.5, 2.0)
    x, u, w, cell = O.load_thermal(ncell, ncps, step, ppc, 0.01, 0.1, 3.5, 99, paired)
    nc = int(np.prod(ncell))
    assert x.shape == (nc * ppc, 3) and np.all(w == 3.5)
    # every particle inside its cell
    gc = np.stack([cell // (ncell[1] * ncell[2]), (cell // ncell[2]) % ncell[1], cell % ncell[2]], 1)
    lo = gc * np.array(step)
    assert np.all(x >= lo) and np.all(x < lo + np.array(step))
    # ppc particles per cell
    assert np.array_equal(np.bincount(cell, minlength=nc), np.full(nc, ppc))
    # supercell-major order: the supercell of every particle is non-decreasing
    nsc = [a // b for a, b in zip(ncell, ncps)]
    sc = (gc[:, 0] // ncps[0] * nsc[1] + gc[:, 1] // ncps[1]) * nsc[2] + gc[:, 2] // ncps[2]
    assert np.all(np.diff(sc) >= 0)
    # slot order inside a supercell: cells interleaved (paired: two particles of a cell side by side)
    cells_sc = int(np.prod(ncps))
    first = cell[:ppc * cells_sc]
    lc = np.stack([(first // (ncell[1] * ncell[2])) % ncps[0], ((first // ncell[2]) % ncell[1]) % ncps[1],
                   (first % ncell[2]) % ncps[2]], 1)
    lin = (lc[:, 0] * ncps[1] + lc[:, 1]) * ncps[2] + lc[:, 2]
    if not paired:
        assert np.array_equal(lin, np.tile(np.arange(cells_sc), ppc))
    else:
        npair = (ppc // 2) * 2 * cells_sc
        expect = np.repeat(np.tile(np.arange(cells_sc), ppc // 2), 2)
        assert np.array_equal(lin[:npair], expect)
        if ppc % 2:
            assert np.array_equal(lin[npair:], np.arange(cells_sc))


def test_loader_momenta_are_maxwellian_with_two_streams():
    ncell, ncps, ppc = (16, 16, 16), (4, 4, 4), 16
    uth, drift = 0.02, 0.1
    x, u, w, cell = O.load_thermal(ncell, ncps, (1.0, 1.0, 1.0), ppc, uth, drift, 1.0, 7, False)
    # particle id parity decides the stream; ids are cell * ppc + k and k = slot // cells (unpaired)
    k = np.tile(np.repeat(np.arange(ppc), 64), (16 // 4) ** 3)
    ids = cell.astype(np.int64) * ppc + k
    even = ids % 2 == 0
    dev = u.copy()
    dev[even, 0] -= drift
    dev[~even, 0] += drift
    n = dev.shape[0]
    for d in range(3):
        assert abs(dev[:, d].mean()) < 4 * uth / math.sqrt(n)
        assert dev[:, d].std() == pytest.approx(uth, rel=4 * math.sqrt(2.0 / n))
    # Gaussian fourth moment (kurtosis 3) and the fraction inside one sigma
    z = dev / uth
    assert (z ** 4).mean() == pytest.approx(3.0, abs=0.08)
    assert (np.abs(z) < 1.0).mean() == pytest.approx(math.erf(1 / math.sqrt(2)), abs=0.005)
    # same seed -> same particles; another seed -> others
    x2, u2, _, _ = O.load_thermal(ncell, ncps, (1.0, 1.0, 1.0), ppc, uth, drift, 1.0, 7, False)
    assert np.array_equal(x, x2) and np.array_equal(u, u2)
    x3, _, _, _ = O.load_thermal(ncell, ncps, (1.0, 1.0, 1.0), ppc, uth, drift, 1.0, 8, False)
    assert not np.array_equal(x, x3)
