"""Edge cases of the periodic step, HIP path against the oracle:

- an ensemble with no particles (the step reduces to the Yee updates and the ghost exchange);
- ragged occupancy: one crowded supercell, one nearly empty, the rest empty, with particles
  streaming into empty supercells across the periodic edges;
- a supercell at the 65,536-particle limit (no slack), and one particle past it.

FP64 at the plain path's bar: supercell membership and particle order bit-exact, positions,
momenta and fields within 1e-12 relative. FP32 split positions at the split form's bounds
(x 2e-6 of the box, u 2e-5, E/B/J 2e-4 relative), particles matched through unique weights.
"""
import numpy as np
import pytest

import oracle as O
import pica_amd as P
from helpers import courant_dt, plasma, random_fields, rel_err

pytestmark = pytest.mark.gpu

TOL32 = {"x": 2e-6, "u": 2e-5, "E": 2e-4, "B": 2e-4, "J": 2e-4}


def _say(msg):
    print("[case] " + msg, flush=True)


def _fp32_exact(x, u):
    """positions with FP32 in-cell offsets and FP32 momenta: the values the split form stores"""
    c = np.floor(x)
    return c + (x - c).astype(np.float32).astype(np.float64), u.astype(np.float32).astype(np.float64)


def _clustered(rng, n, lo, hi, uth, drift):
    """n particles uniform in the box [lo, hi), Maxwellian u with +-drift in x"""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    x = lo + rng.random((n, 3)) * (hi - lo)
    u = rng.standard_normal((n, 3)) * uth
    u[:, 0] += np.where(np.arange(n) % 2 == 0, drift, -drift)
    return x, u


def _unique_weights(w0, n):
    return (w0 * (1.0 + np.arange(n) * 2.0 ** -20)).astype(np.float32).astype(np.float64)


def _run_pair(ctx, prec, ncell, ncps, x, u, w, nsteps, shape, dep, fields=None, slack=None):
    """the same input through the oracle (FP64) and the device (prec); returns both sides"""
    dt = courant_dt()
    og = O.Grid.periodic(ncell)
    if fields is not None:
        random_fields(np.random.default_rng(fields), og, 3e-3, 3e-3)
    g = P.YeeGrid.periodic(ctx, ncell, prec=prec)
    g.set_all(og.f)
    oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps, prec=prec)
    if slack is not None:
        ens.set_slack(*slack)
    if w.size:
        ens.add(x, u, w)
    for _ in range(nsteps):
        O.step(og, oe, dt, shape, dep)
    P.run(ens, g, dt, nsteps, shape, dep)
    return og, oe, g, ens


def _check_fp64(og, oe, g, ens):
    gx, gu, gw, goff = ens.particles()
    assert np.array_equal(goff, oe.off), "supercell membership differs"
    assert np.array_equal(gw, oe.w), "particle order differs"
    if oe.np:
        assert rel_err(gx, oe.positions()) < 1e-12
        assert rel_err(gu, oe.momenta()) < 1e-12
    f = g.get_all()
    for c in range(9):
        if np.abs(og.f[c]).max() > 0:
            assert rel_err(f[c], og.f[c]) < 1e-12, "component %d" % c
        else:
            assert np.abs(f[c]).max() == 0, "component %d" % c


def _check_fp32(og, oe, g, ens, ncell):
    gx, gu, gw, goff = ens.particles()
    assert gx.shape[0] == oe.np
    errs = {}
    for name, comps in (("E", (P.EX, P.EY, P.EZ)), ("B", (P.BX, P.BY, P.BZ)), ("J", (P.JX, P.JY, P.JZ))):
        errs[name] = max(rel_err(g.get(c), og.f[c]) if np.abs(og.f[c]).max() > 0 else float(np.abs(g.get(c)).max())
                         for c in comps)
    if oe.np:
        ia, ib = np.argsort(gw, kind="stable"), np.argsort(oe.w, kind="stable")
        assert np.array_equal(gw[ia], oe.w[ib]), "particle sets differ"
        L = np.array(ncell, np.float64)
        dx = gx[ia] - oe.positions()[ib]
        dx -= L * np.round(dx / L)
        errs["x"] = float(np.abs(dx).max() / L.max())
        errs["u"] = rel_err(gu[ia], oe.momenta()[ib])
    print("FP32 errors vs the FP64 oracle: %s" % errs, flush=True)
    for k, v in errs.items():
        assert v < TOL32[k], "%s error %.3g above %.3g" % (k, v, TOL32[k])


@pytest.mark.parametrize("prec", [P.F64, P.F32])
def test_empty_ensemble_steps_the_fields(ctx, prec):
    _say("empty ensemble prec=%d" % prec)
    ncell = (8, 8, 12)
    none = np.zeros((0, 3))
    og, oe, g, ens = _run_pair(ctx, prec, ncell, (4, 4, 4), none, none, np.zeros(0), 5, P.TSC, P.ESIRKEPOV,
                               fields=3)
    x, u, w, off = ens.particles()
    assert x.shape == (0, 3) and w.size == 0 and np.all(off == 0)
    for c in (P.JX, P.JY, P.JZ):
        assert np.abs(g.get(c)).max() == 0
    if prec == P.F64:
        _check_fp64(og, oe, g, ens)
    else:
        _check_fp32(og, oe, g, ens, ncell)


@pytest.mark.parametrize("prec", [P.F64, P.F32])
@pytest.mark.parametrize("shape,dep", [(P.CIC, P.DIRECT), (P.TSC, P.ESIRKEPOV)])
def test_ragged_supercells(ctx, prec, shape, dep):
    """3,000 particles in supercell (0,0,0), 5 in (3,1,1), 14 empty supercells; the drift
    carries particles through the low and high x edges into empty supercells"""
    _say("ragged supercells prec=%d shape=%d dep=%d" % (prec, shape, dep))
    rng = np.random.default_rng(77 + 2 * shape + dep)
    ncell = (16, 8, 8)
    xa, ua = _clustered(rng, 3000, (0, 0, 0), (4, 4, 4), 0.05, 0.3)
    xb, ub = _clustered(rng, 5, (12, 4, 4), (16, 8, 8), 0.05, 0.3)
    x, u = _fp32_exact(np.concatenate([xa, xb]), np.concatenate([ua, ub]))
    _, w0 = plasma(courant_dt(), 64)
    w = _unique_weights(w0, x.shape[0])
    og, oe, g, ens = _run_pair(ctx, prec, ncell, (4, 4, 4), x, u, w, 12, shape, dep)
    counts = np.diff(oe.off)
    print("supercell counts after 12 steps: %s" % counts.tolist(), flush=True)
    assert (counts > 0).sum() > 2, "no particle reached an empty supercell"
    if prec == P.F64:
        _check_fp64(og, oe, g, ens)
    else:
        _check_fp32(og, oe, g, ens, ncell)


@pytest.mark.parametrize("prec", [P.F64, P.F32])
def test_supercell_at_the_particle_limit(ctx, prec):
    """65,536 particles in one 4^3 supercell with no slack (capacity exactly at the limit),
    3 TSC/Esirkepov steps; one particle more is refused with PK_E_CAPACITY"""
    _say("65536-particle supercell prec=%d" % prec)
    rng = np.random.default_rng(91)
    ncell = (8, 8, 8)
    x, u = _clustered(rng, 65536, (0, 0, 0), (4, 4, 4), 0.05, 0.1)
    x, u = _fp32_exact(x, u)
    _, w0 = plasma(courant_dt(), 1024)
    w = _unique_weights(w0, x.shape[0])
    og, oe, g, ens = _run_pair(ctx, prec, ncell, (4, 4, 4), x, u, w, 3, P.TSC, P.ESIRKEPOV, slack=(0.0, 0))
    if prec == P.F64:
        _check_fp64(og, oe, g, ens)
    else:
        _check_fp32(og, oe, g, ens, ncell)
    over = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), (4, 4, 4), prec=prec)
    over.set_slack(0.0, 0)
    xo = np.concatenate([x, [[1.5, 1.5, 1.5]]])
    uo = np.concatenate([u, [[0.0, 0.0, 0.0]]])
    with pytest.raises(P.PicaError) as ei:
        over.add(xo, uo, np.ones(xo.shape[0]) * w0)
    assert ei.value.code == P.PK_E_CAPACITY
    over.close()


@pytest.mark.parametrize("prec", [P.F64, P.F32])
def test_failed_resort_rebuild_refuses_later_calls(ctx, prec):
    """a re-sort whose rebuild fails (one supercell would pass 65,536 particles) reports
    PK_E_CAPACITY; the ensemble then refuses every step and read with PK_E_STATE until a new
    upload, after which it steps normally"""
    _say("failed re-sort rebuild prec=%d" % prec)
    rng = np.random.default_rng(92)
    ncell = (8, 8, 8)
    xa, ua = _clustered(rng, 65536, (0.1, 0.1, 0.1), (3.9, 3.9, 3.9), 0.0, 0.0)
    xb = np.array([[4.02, 1.5, 1.5]])                   # supercell (1, 0, 0), about to cross into (0, 0, 0)
    ub = np.array([[-0.5, 0.0, 0.0]])
    x, u = _fp32_exact(np.concatenate([xa, xb]), np.concatenate([ua, ub]))
    _, w0 = plasma(courant_dt(), 1024)
    w = np.full(x.shape[0], w0)
    g = P.YeeGrid.periodic(ctx, ncell, prec=prec)
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), (4, 4, 4), prec=prec)
    ens.set_slack(0.0, 0)
    ens.add(x, u, w)
    dt = courant_dt()
    with pytest.raises(P.PicaError) as ei:
        P.step(ens, g, dt, P.CIC, P.DIRECT)
    assert ei.value.code == P.PK_E_CAPACITY
    with pytest.raises(P.PicaError) as ei:
        P.step(ens, g, dt, P.CIC, P.DIRECT)
    assert ei.value.code == P.PK_E_STATE
    with pytest.raises(P.PicaError) as ei:
        ens.particles()
    assert ei.value.code == P.PK_E_STATE
    # fresh contents make it usable again
    ens.add(x[-1000:], u[-1000:], w[-1000:])
    g2 = P.YeeGrid.periodic(ctx, ncell, prec=prec)
    P.run(ens, g2, dt, 2, P.CIC, P.DIRECT)
    gx, gu, gw, goff = ens.particles()
    assert gx.shape[0] == 1000
    ens.close()
