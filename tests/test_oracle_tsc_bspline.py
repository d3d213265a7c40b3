"""The oracle's TSC weights against the quadratic B-spline evaluated at each node's distance.

pica has no second-order form factor (its README claims one; the code has CIC only), so the
oracle's TSC (``or_tsc_node`` / ``or_tsc_w`` in ``oracle/pica_oracle.c``) is pinned here by the
closed form of the quadratic B-spline (Hockney & Eastwood):

    W(r) = 3/4 - r^2            |r| <= 1/2
           (3/2 - |r|)^2 / 2    1/2 < |r| < 3/2
           0                    otherwise

summed by brute force over *every* node of the grid, with node positions taken from the Yee
staggering (grid origin -2h: array index a of a component sits at a - 2 + shift).  A node
anchored one cell too low (``floor(xi)`` instead of ``floor(xi + 1/2)``) still has unit sum,
the right first moment and reproduces linear fields, but gives other nodes and negative
weights; it fails every test below.  Offsets sweep [0, 1) including 0, 1/2 - ulp, 1/2,
1/2 + ulp and 1 - ulp, for collocated and staggered nodes alike.

The Esirkepov current is checked the same way: S0 and S1 are the B-spline (or CIC hat) of the
old and new positions at every cell centre, W_x = DS_x (S0_y S0_z + DS_y S0_z / 2 + S0_y DS_z / 2
+ DS_y DS_z / 3) (Esirkepov, CPC 135 (2001) 144, eq. 38) and J_x on x face f is
-q w h_x / dt times the prefix sum of W_x over the cell centres below f.
"""
import math

import numpy as np
import pytest

import oracle as O

# Yee staggering of E (= J) and B: shift of each component's nodes in cells (YeeGrid.h:11-14)
SHIFT = {O.EX: (0, .5, .5), O.EY: (.5, 0, .5), O.EZ: (.5, .5, 0), O.BX: (.5, 0, 0), O.BY: (0, .5, 0),
         O.BZ: (0, 0, .5)}
G = O.GHOST


def w_tsc(r):
    a = np.abs(r)
    return np.where(a <= 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def w_cic(r):
    return np.maximum(0.0, 1.0 - np.abs(r))


def _offsets():
    half_lo = np.nextafter(0.5, 0.0)
    half_hi = np.nextafter(0.5, 1.0)
    one_lo = np.nextafter(1.0, 0.0)
    special = [0.0, 1e-12, 0.25, half_lo, 0.5, half_hi, 0.75, 1.0 - 1e-12, one_lo]
    rng = np.random.default_rng(31)
    return np.array(special + list(rng.random(23)))


def _positions(cell0, offs, rng):
    """positions cell0 + (cell, offset) per dimension, every offset used in every dimension"""
    n = len(offs)
    pos = np.empty((n, 3))
    for d in range(3):
        pos[:, d] = cell0[d] + rng.integers(0, 4, n) + offs[rng.permutation(n)]
    return pos


def _weights_1d(shape_w, p, s, shift):
    """weights of every array index 0..s-1 along one dimension for positions p"""
    nodes = np.arange(s) - G + shift
    return shape_w(p[:, None] - nodes[None, :])


def test_tsc_bspline_closed_form_values():
    """W is the quadratic B-spline: its three nonzero values at offset d are the textbook
    1/2 (1/2 - d)^2, 3/4 - d^2, 1/2 (1/2 + d)^2, it is continuous at |r| = 1/2 and 3/2, and
    it is non-negative"""
    for d in np.linspace(-0.5, 0.5, 11):
        assert math.isclose(w_tsc(d - 1.0), 0.5 * (0.5 + d) ** 2, abs_tol=1e-15)
        assert math.isclose(w_tsc(d), 0.75 - d * d, abs_tol=1e-15)
        assert math.isclose(w_tsc(d + 1.0), 0.5 * (0.5 - d) ** 2, abs_tol=1e-15)
    assert math.isclose(w_tsc(np.nextafter(0.5, 0)), w_tsc(np.nextafter(0.5, 1)), abs_tol=1e-15)
    assert w_tsc(1.5) == 0.0 and w_tsc(np.nextafter(1.5, 0)) < 1e-30
    assert np.all(w_tsc(np.linspace(-3, 3, 1001)) >= 0)


def test_tsc_interpolation_weights_are_the_bspline():
    """interpolated value = sum over all nodes of W(dx) W(dy) W(dz) F(node), F random, for every
    E and B component (collocated and staggered dimensions), over the swept offsets"""
    rng = np.random.default_rng(32)
    ncell, scmin, ncps = (12, 12, 12), (4.0, 4.0, 4.0), (4, 4, 4)
    g = O.Grid.periodic(ncell)
    g.f[:6] = rng.standard_normal(g.f[:6].shape)
    offs = _offsets()
    pos = _positions(scmin, offs, rng)
    e, b = O.interp_sc(g, scmin, ncps, O.TSC, pos)
    got = np.concatenate([e, b], 1)
    for c, sh in SHIFT.items():
        wx, wy, wz = (_weights_1d(w_tsc, pos[:, d], g.s[d], sh[d]) for d in range(3))
        expect = np.einsum("pa,pb,pc,abc->p", wx, wy, wz, g.f[c])
        np.testing.assert_allclose(got[:, c], expect, rtol=0, atol=1e-13 * np.abs(g.f[c]).max())


def test_tsc_interpolation_weights_are_nonnegative_and_sum_to_one():
    """each of the 27 nodes of a particle gets exactly W(dx) W(dy) W(dz) >= 0 (unit impulses)"""
    rng = np.random.default_rng(33)
    ncell, scmin, ncps = (12, 12, 12), (4.0, 4.0, 4.0), (4, 4, 4)
    offs = _offsets()
    pos = _positions(scmin, offs, rng)
    for c in (O.EX, O.BX):                       # one component with each staggering mix per dimension
        sh = SHIFT[c]
        total = np.zeros(len(pos))
        # every node any of the positions (in [4, 8)) can reach: array indices 5..11 per dimension
        for a in range(4, 13):
            for bb in range(4, 13):
                for cc in range(4, 13):
                    g = O.Grid.periodic(ncell)
                    g.f[c][a, bb, cc] = 1.0
                    e, b = O.interp_sc(g, scmin, ncps, O.TSC, pos)
                    val = np.concatenate([e, b], 1)[:, c]
                    expect = (w_tsc(pos[:, 0] - (a - G + sh[0])) * w_tsc(pos[:, 1] - (bb - G + sh[1])) *
                              w_tsc(pos[:, 2] - (cc - G + sh[2])))
                    assert np.all(val >= 0.0)
                    np.testing.assert_allclose(val, expect, rtol=0, atol=1e-15)
                    total += val
        np.testing.assert_allclose(total, 1.0, rtol=0, atol=1e-14)


def test_tsc_rho_is_the_bspline():
    """rho of one particle at every cell centre (a - 2 + 1/2) is q w W(dx) W(dy) W(dz) / V >= 0"""
    rng = np.random.default_rng(34)
    ncell = (12, 12, 12)
    offs = _offsets()
    pos = _positions((4.0, 4.0, 4.0), offs, rng)
    for p in pos:
        e = O.Ensemble(ncell, (4, 4, 4), (1.0, 1.0, 1.0), p[None, :], np.zeros((1, 3)), np.array([1.0]),
                       q=1.0, m=1.0)
        g = O.Grid.periodic(ncell)
        rho = O.deposit_rho(g, e, O.TSC)
        wx, wy, wz = (_weights_1d(w_tsc, p[None, d], g.s[d], 0.5)[0] for d in range(3))
        expect = np.einsum("a,b,c->abc", wx, wy, wz)
        phys = (slice(G, -G),) * 3
        # the deposit is folded periodically; the particles sit well inside, so no image overlaps
        assert np.all(rho[phys] >= 0.0)
        np.testing.assert_allclose(rho[phys], expect[phys], rtol=0, atol=1e-15)


def test_tsc_direct_deposit_is_the_bspline():
    """direct TSC deposition puts cur_c W(dx) W(dy) W(dz) / V on J_c's staggered nodes"""
    rng = np.random.default_rng(35)
    ncell, scmin, ncps = (12, 12, 12), (4.0, 4.0, 4.0), (4, 4, 4)
    offs = _offsets()
    pos = _positions(scmin, offs, rng)
    for p in pos:
        g = O.Grid.periodic(ncell)
        cur = np.array([[1.0, 2.0, -3.0]])
        O.deposit_direct_sc(g, scmin, ncps, O.TSC, p[None, :], cur)
        for k, c in enumerate((O.JX, O.JY, O.JZ)):
            sh = SHIFT[O.EX + k]
            wx, wy, wz = (_weights_1d(w_tsc, p[None, d], g.s[d], sh[d])[0] for d in range(3))
            expect = cur[0, k] * np.einsum("a,b,c->abc", wx, wy, wz)
            assert np.all(g.f[c] * np.sign(cur[0, k]) >= 0.0)
            np.testing.assert_allclose(g.f[c], expect, rtol=0, atol=1e-14)


def _esirkepov_brute(shape_w, s, p0, p1, qw, dt, step=(1.0, 1.0, 1.0)):
    """J of one particle moving p0 -> p1 from the shapes at every cell centre (eq. 38 of Esirkepov)"""
    S0 = [_weights_1d(shape_w, np.array([p0[d]]), s[d], 0.5)[0] for d in range(3)]
    S1 = [_weights_1d(shape_w, np.array([p1[d]]), s[d], 0.5)[0] for d in range(3)]
    DS = [S1[d] - S0[d] for d in range(3)]
    out = []
    for comp in range(3):
        d1, d2 = (comp + 1) % 3, (comp + 2) % 3
        T = (np.multiply.outer(S0[d1], S0[d2]) + 0.5 * np.multiply.outer(DS[d1], S0[d2]) +
             0.5 * np.multiply.outer(S0[d1], DS[d2]) + np.multiply.outer(DS[d1], DS[d2]) / 3.0)
        W = np.multiply.outer(DS[comp], T)                  # axes (comp, d1, d2)
        J = np.zeros_like(W)
        J[1:] = np.cumsum(W, axis=0)[:-1]                  # face f collects the centres below it
        J *= -qw * step[comp] / dt
        # back to (x, y, z) axis order
        order = [comp, d1, d2]
        out.append(np.transpose(J, np.argsort(order)))
    return out


@pytest.mark.parametrize("shape,shape_w", [(O.TSC, w_tsc), (O.CIC, w_cic)])
def test_esirkepov_current_from_bspline_shapes(shape, shape_w):
    """Esirkepov J of one particle equals the brute-force sum built from the closed-form
    shapes of its old and new positions, for moves that cross and do not cross a node, from
    the swept offsets"""
    rng = np.random.default_rng(36)
    ncell, scmin, ncps = (12, 12, 12), (4.0, 4.0, 4.0), (4, 4, 4)
    offs = _offsets()
    pos0 = _positions(scmin, offs, rng)
    dt = 0.7
    for i, p0 in enumerate(pos0):
        mv = (rng.random(3) - 0.5) * (0.98 if i % 2 else 0.2)
        if i % 5 == 0:
            mv[:] = 0.0
            mv[i % 3] = 0.45 * (1 if i % 2 else -1)        # along one axis
        p1 = p0 + mv
        g = O.Grid.periodic(ncell)
        O.deposit_esirkepov_sc(g, scmin, ncps, shape, p0[None, :], p1[None, :], np.array([1.3]), dt)
        expect = _esirkepov_brute(shape_w, g.s, p0, p1, 1.3, dt)
        for k, c in enumerate((O.JX, O.JY, O.JZ)):
            scale = max(np.abs(expect[k]).max(), 1e-300)
            np.testing.assert_allclose(g.f[c], expect[k], rtol=0, atol=1e-13 * max(scale, 1.0))


def test_lower_node_anchoring_would_fail():
    """the brute-force check is sharp: weights anchored one node low (floor(xi) instead of
    floor(xi + 1/2)) differ from the B-spline by up to about 1/4 for offsets in [0, 1/2)"""
    d = np.linspace(0.0, 0.49, 50)
    # a mis-anchored TSC uses the nearest-node formula with d' = d (not d - 1) on nodes c-1, c, c+1
    wrong = [0.5 * (0.5 - d) ** 2, 0.75 - d * d, 0.5 * (0.5 + d) ** 2]
    # true weights on nodes c-1, c, c+1 for a particle at c + d (d in [0, 1/2)): W(d + 1), W(d), W(d - 1)
    right = [w_tsc(d + 1.0), w_tsc(d), w_tsc(d - 1.0)]
    assert np.allclose(wrong, right)                     # on [0, 1/2) the anchoring coincides
    d = np.linspace(0.51, 0.99, 50)
    low = [0.5 * (0.5 - d) ** 2, 0.75 - d * d, 0.5 * (0.5 + d) ** 2]   # floor(xi) keeps nodes c-1, c, c+1
    right = [w_tsc(d + 1.0), w_tsc(d), w_tsc(d - 1.0)]
    assert np.abs(np.array(low) - np.array(right)).max() > 0.2
    assert np.array(low).min() < 0.0
