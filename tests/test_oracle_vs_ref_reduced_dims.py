"""The reduced-dimension path against pica's own 1D and 2D templates (oracle/_ref).

The library runs a 1D or 2D problem as the 3D grid with edges of one cell in the reduced
dimensions (every periodic ghost of such an edge images that cell), and the oracle restates that
3D path.  Here the oracle's 3D functions on such grids are compared with pica's reduced-dimension
specialisations compiled from its headers:

- YeeSolver::updateE / updateB for YeeGrid<One> and YeeGrid<Two> (YeeSolver.h:41-128): the curl
  terms along a one-cell edge are differences of identical ghost images, exactly zero, so the
  updates are bit-identical away from pica's open-edge rows;
- FieldInterpolatorCIC<YeeGrid<One>> (FieldInterpolatorYeeGrid.h:51-83) and CurrentDepositorCIC
  for YeeGrid<One> and YeeGrid<Two> (CurrentDepositorYeeGrid.h:52-122): the 3D weights along a
  reduced dimension sum to one over identical values, so they agree to rounding (1e-15).

Two of pica's 2D specialisations cannot serve as a reference:
- FieldInterpolatorCIC<YeeGrid<Two>> does not compile (its get() reads .z of a Vector2,
  FieldInterpolatorYeeGrid.h:108-118);
- CurrentDepositorCIC<YeeGrid<Two>>::deposit passes each J component to its helper by value
  (CurrentDepositorYeeGrid.h:114), so it adds into a temporary copy and leaves the grid unchanged
  (asserted below).  The 2D deposit is checked against pica's 3D depositor on the one-cell-edge
  grid instead.
"""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.skipif(O.ref() is None, reason="pica sources not present (oracle/_ref not built)")

G = O.GHOST
_dp = ctypes.POINTER(ctypes.c_double)


def _comps(arrs):
    keep = [np.ascontiguousarray(a) for a in arrs]
    return keep, (_dp * 9)(*[a.ctypes.data_as(_dp) for a in keep])


def _oracle_grid(ncell, step, rng):
    g = O.Grid.periodic(ncell, step)
    g.f[...] = rng.standard_normal(g.f.shape)
    for c in range(9):
        O.ghost_fill(g.f[c])
    return g


def _plane(g, dims):
    """the reduced-dimension arrays of a one-cell-edge 3D grid: the physical cell's line / plane"""
    if dims == 1:
        return [np.ascontiguousarray(g.f[c][:, G, G]) for c in range(9)]
    return [np.ascontiguousarray(g.f[c][:, :, G]) for c in range(9)]


@pytest.mark.parametrize("dims", [1, 2])
@pytest.mark.parametrize("op", ["e", "b"])
def test_yee_reduced_dims_bit_identical_to_pica(dims, op):
    rng = np.random.default_rng(50 + dims)
    ncell = (12, 1, 1) if dims == 1 else (12, 9, 1)
    step = (1.0, 0.5, 2.0)
    dt = 0.4 / O.C
    g = _oracle_grid(ncell, step, rng)
    keep, ptrs = _comps(_plane(g, dims))
    s = g.s[:dims]
    fn = O.ref().ref_yee_update_1d if dims == 1 else O.ref().ref_yee_update_2d
    fn((ctypes.c_int * dims)(*s), (ctypes.c_double * dims)(*g.origin[:dims]), (ctypes.c_double * dims)(*step[:dims]),
       ptrs, ctypes.c_double(dt), 0 if op == "e" else 1)
    (O.yee_update_e if op == "e" else O.yee_update_b)(g, dt)
    got = _plane(g, dims)
    # pica's open-edge rows differ from the periodic 3D grid's; compare the rows both update the
    # same way: updateE [0, s-1), updateB [1, s) in every kept dimension
    sl = tuple(slice(0, n - 1) if op == "e" else slice(1, n) for n in s)
    comps = (O.EX, O.EY, O.EZ) if op == "e" else (O.BX, O.BY, O.BZ)
    for c in comps:
        assert np.array_equal(got[c][sl], keep[c][sl]), "component %d differs" % c
    # the update did something
    assert any(not np.array_equal(got[c], _plane(_oracle_grid(ncell, step, np.random.default_rng(50 + dims)), dims)[c])
               for c in comps)


def test_interp_1d_matches_pica():
    rng = np.random.default_rng(60)
    ncell, step = (12, 1, 1), (1.0, 0.5, 2.0)
    g = _oracle_grid(ncell, step, rng)
    n = 200
    x = 1.0 + rng.random(n) * 10.0
    pos3 = np.stack([x, rng.random(n) * step[1], rng.random(n) * step[2]], 1)
    e3, b3 = O.interp_cic(g, pos3)
    keep, ptrs = _comps(_plane(g, 1))
    e1, b1 = np.zeros((n, 3)), np.zeros((n, 3))
    O.ref().ref_interp_cic_1d((ctypes.c_int * 1)(g.s[0]), (ctypes.c_double * 1)(g.origin[0]),
                              (ctypes.c_double * 1)(step[0]), ptrs, ctypes.c_long(n),
                              np.ascontiguousarray(x).ctypes.data_as(_dp), e1.ctypes.data_as(_dp),
                              b1.ctypes.data_as(_dp))
    scale = np.abs(g.f[:6]).max()
    np.testing.assert_allclose(e3, e1, rtol=0, atol=2e-15 * scale)
    np.testing.assert_allclose(b3, b1, rtol=0, atol=2e-15 * scale)


def _ref_deposit_3d(g, pos3, cur):
    xs = [np.ascontiguousarray(pos3[:, d]) for d in range(3)]
    O.ref().ref_deposit_cic((ctypes.c_int * 3)(*g.s), (ctypes.c_double * 3)(*g.origin), (ctypes.c_double * 3)(*g.step),
                            g.comps_ptrs(), ctypes.c_long(pos3.shape[0]), *[x.ctypes.data_as(_dp) for x in xs],
                            np.ascontiguousarray(cur).ctypes.data_as(_dp))


@pytest.mark.parametrize("dims", [1, 2])
def test_deposit_reduced_dims_matches_pica(dims):
    rng = np.random.default_rng(70 + dims)
    ncell = (12, 1, 1) if dims == 1 else (12, 9, 1)
    step = (1.0, 0.5, 2.0)
    n = 300
    lo = np.array([2.0, 2.0, 0.0])[:dims] * np.array(step[:dims])
    hi = (np.array(ncell[:dims]) - 2.0) * np.array(step[:dims])
    pd = lo + rng.random((n, dims)) * (hi - lo)                        # away from the x / y edges
    pos3 = np.zeros((n, 3))
    pos3[:, :dims] = pd
    for d in range(dims, 3):
        pos3[:, d] = rng.random(n) * step[d]
    cur = rng.standard_normal((n, 3))
    g = O.Grid.periodic(ncell, step)
    O.deposit_cic(g, pos3, cur)
    for c in (O.JX, O.JY, O.JZ):
        O.ghost_fold(g.f[c])
    g2 = O.Grid.periodic(ncell, step)
    keep, ptrs = _comps(_plane(g2, dims))
    fn = O.ref().ref_deposit_cic_1d if dims == 1 else O.ref().ref_deposit_cic_2d
    fn((ctypes.c_int * dims)(*g.s[:dims]), (ctypes.c_double * dims)(*g.origin[:dims]),
       (ctypes.c_double * dims)(*step[:dims]), ptrs, ctypes.c_long(n),
       np.ascontiguousarray(pd).ctypes.data_as(_dp), np.ascontiguousarray(cur).ctypes.data_as(_dp))
    got = _plane(g, dims)
    phys = tuple(slice(G, G + ncell[d]) for d in range(dims))
    if dims == 2:
        # pica's 2D depositor leaves the grid unchanged (J taken by value); use its 3D depositor
        assert all(np.abs(keep[c]).max() == 0.0 for c in (O.JX, O.JY, O.JZ))
        g3 = O.Grid.periodic(ncell, step)
        _ref_deposit_3d(g3, pos3, cur)
        for c in (O.JX, O.JY, O.JZ):
            O.ghost_fold(g3.f[c])
        keep = _plane(g3, dims)
    for c in (O.JX, O.JY, O.JZ):
        scale = np.abs(keep[c]).max()
        assert scale > 0
        np.testing.assert_allclose(got[c][phys], keep[c][phys], rtol=0, atol=1e-15 * scale * 4)
