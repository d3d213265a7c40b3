"""Closed forms and invariants that pin the oracle where pica has no counterpart (TSC,
Esirkepov, periodic ghosts, re-sort, rho) and re-check pica's own semantics independently
(Boris rotation angle, Yee dispersion, known answers of pica's tests).  Each check is chosen so
that a plausible slip (a dropped term, a wrong sign, index or staggering) fails it.
"""
import math

import numpy as np
import pytest

import oracle as O


# ------------------------------------------------------------------ known answers of pica's tests
def test_getcellindex_of_cell_centres():
    """YeeGridTests.cpp:126-134: the centre of cell (i,j,k) maps to (i,j,k)"""
    g = O.Grid((5, 4, 3), (1.5, -2.0, 0.25), (0.5, 2.0, 1.0))
    idx = np.array([(i, j, k) for i in range(5) for j in range(4) for k in range(3)])
    pos = np.array(g.origin) + np.array(g.step) * (idx + 0.5)
    assert np.array_equal(O.cell_index_grid(g, pos), idx)


def test_storage_is_z_fastest_and_staggering():
    """Array3d stores (i*ny + j)*nz + k (Array.h:143, UtilityTests.cpp:88-92: index (2,3,1) of a
    (4,6,2) array is 31); a unit value at one node interpolates to 1 at that node's position for
    every component, at the positions YeeGrid.h:11-14 gives (Ex at (i, j+1/2, k+1/2), ...)"""
    a = np.zeros((4, 6, 2))
    a.ravel()[31] = 1.0
    assert a[2, 3, 1] == 1.0
    s, origin, h = (8, 8, 8), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    shift = {O.EX: (0, .5, .5), O.EY: (.5, 0, .5), O.EZ: (.5, .5, 0), O.BX: (.5, 0, 0), O.BY: (0, .5, 0),
             O.BZ: (0, 0, .5)}
    for comp, sh in shift.items():
        g = O.Grid(s, origin, h)
        g.f[comp][3, 4, 2] = 1.0
        pos = np.array([[origin[d] + ((3, 4, 2)[d] + sh[d]) * h[d] for d in range(3)]])
        e, b = O.interp_cic(g, pos)
        val = np.concatenate([e[0], b[0]])
        assert val[comp] == pytest.approx(1.0, abs=1e-15)
        assert np.count_nonzero(np.abs(val) > 1e-15) == 1


# ------------------------------------------------------------------ Boris pusher
def _push(u, e, b, dt=1e-12, n=1):
    x = np.zeros((u.shape[0], 3)) + 5.0
    for _ in range(n):
        x, u = O.boris_push(x, u, e, b, O.E_CHARGE, O.E_MASS, dt)
    return x, u


def test_boris_magnetic_rotation_angle():
    """E = 0: |u| is conserved and u turns by 2 atan(|t|) per step, t = (q dt / 2mc) B / gamma
    (the Boris rotation, independent of pica's arithmetic)"""
    rng = np.random.default_rng(1)
    u0 = rng.standard_normal((50, 3)) * 2.0
    bb = rng.standard_normal((50, 3)) * 3e4
    dt = 1e-12
    x, u1 = _push(u0.copy(), np.zeros((50, 3)), bb, dt)
    assert np.allclose(np.linalg.norm(u1, axis=1), np.linalg.norm(u0, axis=1), rtol=1e-13, atol=0)
    gamma = np.sqrt(1 + (u0 ** 2).sum(1))
    t = np.linalg.norm(bb, axis=1) * abs(O.E_CHARGE) * dt / (2 * O.E_MASS * O.C) / gamma
    bh = bb / np.linalg.norm(bb, axis=1)[:, None]
    up0 = u0 - (u0 * bh).sum(1)[:, None] * bh
    up1 = u1 - (u1 * bh).sum(1)[:, None] * bh
    cosang = (up0 * up1).sum(1) / (np.linalg.norm(up0, axis=1) * np.linalg.norm(up1, axis=1))
    assert np.allclose(np.arccos(np.clip(cosang, -1, 1)), 2 * np.arctan(t), rtol=1e-8, atol=1e-12)
    # parallel component untouched
    assert np.allclose((u1 * bh).sum(1), (u0 * bh).sum(1), rtol=1e-13, atol=1e-15)


def test_boris_electric_kick_and_velocity():
    """B = 0: u+ = u- + 2 (q dt / 2mc) E; x advances by c u / gamma dt with gamma = sqrt(1+u^2)
    (ParticleTests.cpp:37-60 pins gamma this way at 1e-12)"""
    rng = np.random.default_rng(2)
    u0 = rng.standard_normal((40, 3))
    ee = rng.standard_normal((40, 3)) * 1e3
    dt = 1e-12
    x, u1 = _push(u0.copy(), ee, np.zeros((40, 3)), dt)
    k = O.E_CHARGE * dt / (2 * O.E_MASS * O.C)
    assert np.allclose(u1, u0 + 2 * k * ee, rtol=1e-13, atol=1e-15)
    g1 = np.sqrt(1 + (u1 ** 2).sum(1))
    assert np.allclose(x - 5.0, u1 * O.C / g1[:, None] * dt, rtol=1e-12, atol=0)


# ------------------------------------------------------------------ Yee solver
def test_yee_plane_wave_discrete_dispersion():
    """A vacuum plane wave along x with the discrete eigen-frequency
    sin(w dt / 2) / (c dt) = sin(k h / 2) / h stays a plane wave: after 40 steps of B/2, E, B/2,
    Ey = cos(k x - w t) to round-off.  Leapfrog bookkeeping: B at half steps is
    A cos(k x - w (n + 1/2) dt) with A = h sin(w dt/2) / (c dt sin(k h/2)); the integer-time B the
    step starts from is B^0 = B^(1/2) - (dt/2) F(E^0).  A wrong coefficient, sign or staggering fails."""
    N = 32
    h = 1.0
    dt = 0.5 * h / O.C
    k = 2 * math.pi * 3 / N
    w = 2.0 / dt * math.asin(O.C * dt / h * math.sin(k * h / 2))
    A = h * math.sin(w * dt / 2) / (O.C * dt * math.sin(k * h / 2))
    g = O.Grid.periodic((N, 4, 4), (h, h, h))
    G = O.GHOST
    i = np.arange(g.s[0]) - G
    xe = (i + 0.5) * h      # Ey at (i + 1/2, j, k + 1/2)
    xb = i * h              # Bz at (i, j, k + 1/2)
    g.f[O.EY][...] = np.cos(k * xe)[:, None, None]
    b0 = A * np.cos(k * xb - w * dt / 2) - (O.C * dt / h) * np.sin(k * xb) * math.sin(k * h / 2)
    g.f[O.BZ][...] = b0[:, None, None]
    for c in range(6):
        O.ghost_fill(g.f[c])
    nsteps = 40
    for _ in range(nsteps):
        O.field_stage(g, dt)
    ey = g.physical(O.EY)[:, 0, 0]
    expect = np.cos(k * xe[G:G + N] - w * nsteps * dt)
    assert np.max(np.abs(ey - expect)) < 1e-11
    # the continuous dispersion relation would be off by far more than round-off
    assert np.max(np.abs(ey - np.cos(k * xe[G:G + N] - O.C * k * nsteps * dt))) > 1e-4


def test_yee_divergences():
    """periodic: from B = 0, div B stays zero to round-off; div E changes by -4 pi dt div J
    (the discrete divergence of the discrete curl vanishes)"""
    rng = np.random.default_rng(3)
    g = O.Grid.periodic((12, 10, 8))
    G = O.GHOST
    for c in (O.EX, O.EY, O.EZ):
        g.f[c][...] = rng.standard_normal(g.s)
        O.ghost_fill(g.f[c])
    for c in (O.JX, O.JY, O.JZ):
        g.f[c][G:-G, G:-G, G:-G] = rng.standard_normal(tuple(n - 2 * G for n in g.s))   # empty ghosts: fold is a no-op
    dt = 0.4 / O.C
    dive0 = O.physical_nodes(O.divergence(g.f[O.EX], g.f[O.EY], g.f[O.EZ], g.step))
    O.field_stage(g, dt)
    divj = O.physical_nodes(O.divergence(g.f[O.JX], g.f[O.JY], g.f[O.JZ], g.step))
    dive1 = O.physical_nodes(O.divergence(g.f[O.EX], g.f[O.EY], g.f[O.EZ], g.step))
    bx, by, bz = g.f[O.BX], g.f[O.BY], g.f[O.BZ]
    divb = (bx[1:, 1:, 1:] - bx[:-1, 1:, 1:]) + (by[1:, 1:, 1:] - by[1:, :-1, 1:]) + (bz[1:, 1:, 1:] - bz[1:, 1:, :-1])
    divb = divb[G:-G, G:-G, G:-G]
    scale_b = max(np.abs(bx).max(), np.abs(by).max(), np.abs(bz).max())
    assert scale_b > 0 and np.abs(divb).max() < 1e-13 * scale_b
    assert np.allclose(dive1 - dive0, -4 * O.PI * dt * divj, rtol=0, atol=1e-12 * np.abs(dive0).max())


@pytest.mark.parametrize("shape", [(10, 9, 8), (5, 9, 5), (6, 5, 8), (5, 5, 5)])
def test_ghost_fill_and_fold_are_periodic(shape):
    """fill copies the periodic image, fold adds every ghost into its image; edges shorter than
    the ghost width (1D / 2D grids: one physical cell) wrap more than once"""
    rng = np.random.default_rng(4)
    a = rng.standard_normal(shape)
    b = a.copy()
    O.ghost_fill(b)
    G = 2
    N = np.array(a.shape) - 2 * G
    img = [(np.arange(n) - G) % m + G for n, m in zip(a.shape, N)]
    for i in range(a.shape[0]):
        for j in range(a.shape[1]):
            for k in range(a.shape[2]):
                assert b[i, j, k] == a[img[0][i], img[1][j], img[2][k]]
    # fold: each physical value is the sum of every array entry imaging to it (brute force)
    c = rng.standard_normal(shape)
    ref = np.zeros(shape)
    for i in range(a.shape[0]):
        for j in range(a.shape[1]):
            for k in range(a.shape[2]):
                ref[img[0][i], img[1][j], img[2][k]] += c[i, j, k]
    O.ghost_fold(c)
    phys = (slice(G, -G),) * 3
    np.testing.assert_allclose(c[phys], ref[phys], rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------------ TSC / rho / Esirkepov
def _single_particle_ensemble(x, u=(0.0, 0.0, 0.0), ncell=(8, 8, 8), w=1.0):
    return O.Ensemble(ncell, (4, 4, 4), (1.0, 1.0, 1.0), np.array([x]), np.array([u]), np.array([w]), q=1.0, m=1.0)


@pytest.mark.parametrize("shape", [O.CIC, O.TSC])
def test_rho_partition_of_unity_and_first_moment(shape):
    """sum rho V = q w and the rho-weighted mean of the cell centres is the particle position
    (exact for CIC and TSC: partition of unity and linear moment)"""
    rng = np.random.default_rng(5)
    for _ in range(20):
        x = 2.0 + rng.random(3) * 4.0
        e = _single_particle_ensemble(x)
        g = O.Grid.periodic((8, 8, 8))
        rho = O.deposit_rho(g, e, shape)
        ph = rho[2:-2, 2:-2, 2:-2]
        assert math.isclose(ph.sum(), 1.0, rel_tol=1e-14)
        centres = [np.arange(8) + 0.5] * 3
        mom = [(ph.sum(axis=tuple(a for a in range(3) if a != d)) * centres[d]).sum() for d in range(3)]
        assert np.allclose(mom, x, rtol=0, atol=1e-13)


@pytest.mark.parametrize("shape", [O.CIC, O.TSC])
def test_linear_field_is_interpolated_exactly(shape):
    """a field linear in each coordinate is reproduced exactly by CIC and TSC weights"""
    g = O.Grid.periodic((8, 8, 8))
    s = g.s
    ii, jj, kk = np.meshgrid(np.arange(s[0]), np.arange(s[1]), np.arange(s[2]), indexing="ij")
    shift = {O.EX: (0, .5, .5), O.EY: (.5, 0, .5), O.EZ: (.5, .5, 0), O.BX: (.5, 0, 0), O.BY: (0, .5, 0),
             O.BZ: (0, 0, .5)}
    coef = {c: np.array([0.3 * c + 1.0, -0.7, 0.2 + 0.1 * c, 0.5]) for c in shift}
    for c, sh in shift.items():
        X = ii - 2 + sh[0]
        Y = jj - 2 + sh[1]
        Z = kk - 2 + sh[2]
        a = coef[c]
        g.f[c][...] = a[0] + a[1] * X + a[2] * Y + a[3] * Z
    rng = np.random.default_rng(6)
    pos = np.array([4.0, 0.0, 4.0]) + rng.random((50, 3)) * 4.0
    e, b = O.interp_sc(g, (4.0, 0.0, 4.0), (4, 4, 4), shape, pos)
    val = np.concatenate([e, b], 1)
    for c in shift:
        a = coef[c]
        expect = a[0] + a[1] * pos[:, 0] + a[2] * pos[:, 1] + a[3] * pos[:, 2]
        assert np.allclose(val[:, c], expect, rtol=1e-13, atol=1e-13)


def _esirkepov_one(shape, x0, x1, qw=1.0, dt=1.0, scmin=(4.0, 4.0, 4.0)):
    g = O.Grid.periodic((12, 12, 12))
    O.deposit_esirkepov_sc(g, scmin, (4, 4, 4), shape, np.array([x0]), np.array([x1]), np.array([qw]), dt)
    return g


@pytest.mark.parametrize("shape", [O.CIC, O.TSC])
def test_esirkepov_single_particle_continuity(shape):
    """(rho1 - rho0)/dt + div J = 0 at every node to round-off, for random moves < 1/2 cell"""
    rng = np.random.default_rng(7)
    dt = 0.7
    for _ in range(30):
        x0 = 4.0 + rng.random(3) * 4.0
        x1 = x0 + (rng.random(3) - 0.5) * 0.98
        g = _esirkepov_one(shape, x0, x1, 1.0, dt)
        r0 = O.deposit_rho(O.Grid.periodic((12, 12, 12)), _single_particle_ensemble(x0, ncell=(12, 12, 12)), shape)
        r1 = O.deposit_rho(O.Grid.periodic((12, 12, 12)), _single_particle_ensemble(x1, ncell=(12, 12, 12)), shape)
        for c in (O.JX, O.JY, O.JZ):
            O.ghost_fold(g.f[c])
            O.ghost_fill(g.f[c])
        divj = O.divergence(g.f[O.JX], g.f[O.JY], g.f[O.JZ], g.step)
        res = O.physical_nodes(((r1 - r0) / dt)[:-1, :-1, :-1] + divj)
        assert np.abs(res).max() < 1e-14 * max(1.0, np.abs(divj).max())


@pytest.mark.parametrize("shape", [O.CIC, O.TSC])
def test_esirkepov_stationary_and_axis_moves(shape):
    """a particle at rest deposits nothing; a move along x deposits only Jx and
    sum Jx V = q w vx (the total current of a point charge)"""
    g = _esirkepov_one(shape, (5.3, 6.1, 4.7), (5.3, 6.1, 4.7))
    assert np.abs(g.f[O.JX:O.JZ + 1]).max() == 0.0
    dt = 0.5
    g = _esirkepov_one(shape, (5.3, 6.1, 4.7), (5.7, 6.1, 4.7), 2.0, dt)
    assert np.abs(g.f[O.JY]).max() == 0.0 and np.abs(g.f[O.JZ]).max() == 0.0
    assert math.isclose(g.f[O.JX].sum(), 2.0 * 0.4 / dt, rel_tol=1e-13)


def test_esirkepov_cic_equals_villasenor_buneman_within_a_cell():
    """for CIC and a move inside one cell, Esirkepov's current equals the Villasenor-Buneman
    closed form: Jx on the four x-edges = q dx/dt * ((1 - y)(1 - z) + dy dz / 12) etc., with y, z the
    midpoint fractions (here relative to the cell-centre grid)"""
    dt = 1.0
    x0 = np.array([5.60, 6.55, 4.70])
    x1 = np.array([5.90, 6.80, 4.95])
    g = _esirkepov_one(O.CIC, x0, x1, 1.0, dt)
    d = x1 - x0
    # cell-centre coordinates: node a sits at a - 1.5 (array index a, ghost 2), fractions inside
    xi0, xi1 = x0 + 1.5, x1 + 1.5
    base = np.floor(xi0).astype(int)
    assert np.array_equal(base, np.floor(xi1).astype(int))
    mid = (xi0 + xi1) / 2 - base
    # Jx at face index base_x + 1 (between the two x nodes), transverse nodes base_y + (0,1), base_z + (0,1)
    jx = g.f[O.JX]
    fy, fz = mid[1], mid[2]
    expect = {
        (0, 0): d[0] / dt * ((1 - fy) * (1 - fz) + d[1] * d[2] / 12),
        (1, 0): d[0] / dt * (fy * (1 - fz) - d[1] * d[2] / 12),
        (0, 1): d[0] / dt * ((1 - fy) * fz - d[1] * d[2] / 12),
        (1, 1): d[0] / dt * (fy * fz + d[1] * d[2] / 12),
    }
    for (a, b), v in expect.items():
        got = jx[base[0] + 1, base[1] + a, base[2] + b]
        assert math.isclose(got, v, rel_tol=1e-12), (a, b, got, v)


# ------------------------------------------------------------------ ensemble and re-sort
def test_resort_keeps_every_particle_in_its_supercell():
    rng = np.random.default_rng(8)
    ncell, ncps = (12, 8, 8), (4, 4, 2)
    n = 3000
    x = rng.random((n, 3)) * np.array(ncell)
    u = rng.standard_normal((n, 3)) * 0.3
    w = np.arange(n, dtype=np.float64) + 1.0          # unique weights identify particles
    e = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    g = O.Grid.periodic(ncell)
    dt = 0.5 / O.C
    for _ in range(5):
        O.step(g, e, dt, O.TSC, O.ESIRKEPOV)
        sc = O.supercell_index((0, 0, 0), ncell, ncell, ncps, e.positions())
        lin = (sc[:, 0] * e.nsc[1] + sc[:, 1]) * e.nsc[2] + sc[:, 2]
        owner = np.repeat(np.arange(len(e.off) - 1), np.diff(e.off))
        assert np.array_equal(lin, owner)
        assert np.array_equal(np.sort(e.w), w)           # nothing lost or duplicated
        assert np.all(e.positions() >= 0) and np.all(e.positions() < np.array(ncell))
