"""The CPU oracle against pica's own code on fresh random inputs (needs oracle/_ref, built
from the pica sources where they are present; skipped elsewhere -- the committed golden
fixtures cover that case)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.skipif(O.ref() is None, reason="pica reference harness not built here")


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_yee(seed, dtype):
    rng = np.random.default_rng(seed)
    s = tuple(int(v) for v in rng.integers(3, 11, 3))
    step = tuple(float(v) for v in rng.random(3) + 0.5)
    g = O.Grid(s, (-1.0, 0.5, -2.0), step, dtype, rng.standard_normal((9,) + s))
    r = g.copy()
    for kind, dt in (("e", 3e-11), ("b", 1e-11), ("e", 2e-11)):
        (O.yee_update_e if kind == "e" else O.yee_update_b)(g, dtype(dt))
        O.REF.yee(kind, r, dtype(dt))
        assert np.array_equal(g.f, r.f)


@pytest.mark.parametrize("seed", range(3))
def test_interp_deposit_plain_and_supercell(seed):
    rng = np.random.default_rng(100 + seed)
    s = (16, 16, 16)
    g = O.Grid(s, (-2.0, -2.0, -2.0), (1.0, 1.0, 1.0), f=rng.standard_normal((9,) + s))
    pos = rng.random((500, 3)) * 9.0 + 0.5
    e, b = O.interp_cic(g, pos)
    er, br = O.REF.interp_cic(g, pos)
    assert np.array_equal(e, er) and np.array_equal(b, br)
    scmin = (4.0, 8.0, 0.0)
    psc = np.array(scmin) + rng.random((300, 3)) * 4.0
    e, b = O.interp_sc(g, scmin, (4, 4, 4), O.CIC, psc)
    er, br = O.REF.interp_cic(g, psc, scmin, (4, 4, 4))
    assert np.array_equal(e, er) and np.array_equal(b, br)
    cur = rng.standard_normal((500, 3))
    g1, g2 = O.Grid(s, g.origin, g.step), O.Grid(s, g.origin, g.step)
    O.deposit_cic(g1, pos, cur)
    O.REF.deposit_cic(g2, pos, cur)
    assert np.array_equal(g1.f, g2.f)
    pds = np.array(scmin) - 0.5 + rng.random((300, 3)) * 5.0
    g1, g2 = O.Grid(s, g.origin, g.step), O.Grid(s, g.origin, g.step)
    O.deposit_direct_sc(g1, scmin, (4, 4, 4), O.CIC, pds, cur[:300])
    O.REF.deposit_cic(g2, pds, cur[:300], scmin, (4, 4, 4))
    assert np.array_equal(g1.f, g2.f)


def test_supercell_variants_agree_with_plain_for_origin_minus_two_steps():
    """pica's supercell classes ignore the grid origin and assume -2 h; for that origin they match
    the plain classes to rounding (differences of order 1e-15)"""
    rng = np.random.default_rng(7)
    s = (14, 14, 14)
    g = O.Grid(s, (-2.0, -2.0, -2.0), (1.0, 1.0, 1.0), f=rng.standard_normal((9,) + s))
    scmin = (4.0, 4.0, 4.0)
    pos = np.array(scmin) + rng.random((300, 3)) * 4.0
    e1, b1 = O.interp_cic(g, pos)
    e2, b2 = O.interp_sc(g, scmin, (4, 4, 4), O.CIC, pos)
    assert np.abs(e1 - e2).max() < 1e-13 and np.abs(b1 - b2).max() < 1e-13


@pytest.mark.parametrize("seed", range(3))
def test_boris(seed):
    rng = np.random.default_rng(200 + seed)
    n = 300
    x = rng.random((n, 3)) * 5
    u = rng.standard_normal((n, 3)) * 10.0 ** rng.integers(-3, 2, (n, 1))
    e = rng.standard_normal((n, 3)) * 100.0
    b = rng.standard_normal((n, 3)) * 1e3
    w = np.ones(n)
    for dt in (1e-13, 3e-12):
        xo, uo = O.boris_push(x, u, e, b, O.E_CHARGE, O.E_MASS, dt)
        xr, ur = O.REF.boris(x, u, e, b, w, O.E_CHARGE, O.E_MASS, dt)
        assert np.array_equal(xo, xr) and np.array_equal(uo, ur)


def test_indices_including_cell_boundaries():
    rng = np.random.default_rng(300)
    pos = np.concatenate([rng.random((400, 3)) * 30 - 5, np.round(rng.random((100, 3)) * 24) / 2,
                          np.nextafter(np.round(rng.random((50, 3)) * 24), -np.inf)])
    g = O.Grid((30, 30, 30), (-3.0, 0.5, 1.0), (0.5, 1.0, 2.0))
    assert np.array_equal(O.cell_index_grid(g, pos), O.REF.cell_index(g, pos))
    for ncps in ((1, 1, 1), (4, 2, 3)):
        a = O.supercell_index((0.0, 0.0, 0.0), (24.0, 12.0, 36.0), (24, 12, 36), ncps, pos)
        b = O.REF.supercell_index((0.0, 0.0, 0.0), (24.0, 12.0, 36.0), (24, 12, 36), ncps, pos)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("ncell,ncps,ppc,steps", [((8, 8, 8), (4, 4, 4), 4, 8), ((12, 8, 4), (4, 2, 2), 3, 6),
                                                  ((16, 16, 8), (4, 4, 4), 6, 10)])
def test_periodic_run(ncell, ncps, ppc, steps):
    rng = np.random.default_rng(sum(ncell) + steps)
    n = int(np.prod(ncell)) * ppc
    x = rng.random((n, 3)) * np.array(ncell)
    u = rng.standard_normal((n, 3)) * 0.1
    w = rng.random(n) + 1e9
    f = np.zeros((9,) + tuple(c + 4 for c in ncell))
    f[:6] = rng.standard_normal(f[:6].shape) * 1e-3
    for c in range(6):
        O.ghost_fill(f[c])
    dt = 0.5 / O.C
    xr, ur, wr, offr, gr = O.ref_run_cic_direct(ncell, (1.0, 1.0, 1.0), ncps, O.E_CHARGE, O.E_MASS, dt, steps, x, u,
                                                w, f)
    ens = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    g = O.Grid.periodic(ncell)
    g.f[...] = f
    for _ in range(steps):
        O.step(g, ens, dt, O.CIC, O.DIRECT)
    assert np.array_equal(ens.off, offr)
    assert np.array_equal(ens.positions(), xr) and np.array_equal(ens.momenta(), ur) and np.array_equal(ens.w, wr)
    assert np.array_equal(g.f, gr.f)
