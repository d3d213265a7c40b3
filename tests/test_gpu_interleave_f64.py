"""The FP64 ensembles' periodic reorder (PK_OPT_INTERLEAVE_F64) against the oracle.

The reorder changes nothing but the order of the particles inside each supercell: after a step
with it, every supercell holds the oracle's particles in rank-major order (the r-th particle of
every cell, by cell, for r = 0, 1, ..., ranks taken in the oracle's re-sort order), restated here
in numpy; over many steps positions, momenta and fields stay within 1e-12 of the oracle's, with
membership and per-cell counts bit-exact.  Particles carry distinct weights, so the weight column
names every particle.
"""
import numpy as np
import pytest

import oracle as O
import pica_amd as P
from helpers import courant_dt, plasma, rel_err

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case] " + msg, flush=True)


def _two_stream(rng, ncell, ppc, uth=0.01, drift=0.1):
    n = int(np.prod(ncell)) * ppc
    x = rng.random((n, 3)) * np.asarray(ncell, np.float64)
    u = rng.standard_normal((n, 3)) * uth
    u[:, 0] += np.where(np.arange(n) % 2 == 0, drift, -drift)
    return x, u


def _rank_major(w, x, off, ncell, ncps):
    """the weights of each supercell [off[s], off[s + 1]) reordered rank-major by local cell"""
    cell = np.floor(x).astype(np.int64) % np.asarray(ncell)
    nsc = np.asarray(ncell) // np.asarray(ncps)
    out = np.empty_like(w)
    for s in range(len(off) - 1):
        a, b = int(off[s]), int(off[s + 1])
        if b - a < 2:
            out[a:b] = w[a:b]
            continue
        sidx = np.array([s // (nsc[1] * nsc[2]), (s // nsc[2]) % nsc[1], s % nsc[2]])
        loc = cell[a:b] - sidx * np.asarray(ncps)
        lc = (loc[:, 0] * ncps[1] + loc[:, 1]) * ncps[2] + loc[:, 2]
        rank = np.zeros(b - a, np.int64)
        seen = {}
        for i, c in enumerate(lc):
            rank[i] = seen.get(c, 0)
            seen[c] = rank[i] + 1
        order = np.lexsort((lc, rank))
        out[a:b] = w[a:b][order]
    return out


@pytest.mark.parametrize("ncps", [(4, 4, 4), (2, 4, 4)])
def test_reorder_is_rank_major(ctx, ncps):
    _say("FP64 interleave order ncps=%s" % (ncps,))
    ncell, ppc = (16, 16, 16), 8
    dt = courant_dt()
    n0, w0 = plasma(dt, ppc)
    rng = np.random.default_rng(31)
    x, u = _two_stream(rng, ncell, ppc)
    w = w0 * (1.0 + np.arange(x.shape[0]) * 2.0 ** -24)
    ctx.set_option(P.OPT_INTERLEAVE_F64, 1)
    try:
        ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps)
        ens.add(x, u, w)
        oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
        og = O.Grid.periodic(ncell)
        g = P.YeeGrid.periodic(ctx, ncell)
        O.step(og, oe, dt, O.TSC, O.ESIRKEPOV)
        P.step(ens, g, dt, P.TSC, P.ESIRKEPOV)
        gx, gu, gw, goff = ens.particles()
        assert np.array_equal(goff, oe.off), "supercell membership differs"
        want = _rank_major(oe.w, oe.positions(), oe.off, ncell, ncps)
        assert np.array_equal(gw, want), "order is not the rank-major order of the oracle's re-sort"
        # the loader's order is already rank-major when every cell holds the same count: a first
        # reorder of a freshly loaded ensemble leaves it as it is
        fresh = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps)
        fresh.load_thermal(ppc, 0.01, 0.1, w0, 5)
        before = fresh.particles()[2]
        P.step(fresh, P.YeeGrid.periodic(ctx, ncell), dt * 1e-9, P.TSC, P.ESIRKEPOV)
        fx, _, fw, foff = fresh.particles()
        assert np.array_equal(np.sort(fw), np.sort(before))
        assert np.array_equal(fw, _rank_major(fw, fx, foff, ncell, ncps))
    finally:
        ctx.set_option(P.OPT_INTERLEAVE_F64, 0)


def test_reordered_run_keeps_parity(ctx):
    _say("FP64 interleave every 3rd re-sort, 40 steps, TSC/Esirkepov")
    ncell, ncps, ppc = (16, 16, 16), (4, 4, 4), 16
    dt = courant_dt()
    n0, w0 = plasma(dt, ppc)
    rng = np.random.default_rng(32)
    x, u = _two_stream(rng, ncell, ppc)
    w = w0 * (1.0 + np.arange(x.shape[0]) * 2.0 ** -24)
    ctx.set_option(P.OPT_INTERLEAVE_F64, 3)
    try:
        ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps)
        ens.add(x, u, w)
        oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
        og = O.Grid.periodic(ncell)
        g = P.YeeGrid.periodic(ctx, ncell)
        for _ in range(40):
            O.step(og, oe, dt, O.TSC, O.ESIRKEPOV)
        P.run(ens, g, dt, 40, P.TSC, P.ESIRKEPOV)
        gx, gu, gw, goff = ens.particles()
        assert np.array_equal(goff, oe.off), "supercell membership differs"
        counts, _ = ens.cells()
        cell = np.floor(oe.positions()).astype(np.int64) % np.asarray(ncell)
        lin = (cell[:, 0] * ncell[1] + cell[:, 1]) * ncell[2] + cell[:, 2]
        assert np.array_equal(counts.ravel(), np.bincount(lin, minlength=int(np.prod(ncell))))
        # the same particles in each supercell, in another order: match them by weight
        gi, oi = np.argsort(gw, kind="stable"), np.argsort(oe.w, kind="stable")
        assert np.array_equal(gw[gi], oe.w[oi])
        ex = rel_err(gx[gi], oe.positions()[oi])
        eu = rel_err(gu[gi], oe.momenta()[oi])
        f = g.get_all()
        ef = max(rel_err(f[c], og.f[c]) for c in range(9))
        print("reordered FP64 run: x %.3g u %.3g fields %.3g" % (ex, eu, ef), flush=True)
        assert ex < 1e-12 and eu < 1e-12 and ef < 1e-12
    finally:
        ctx.set_option(P.OPT_INTERLEAVE_F64, 0)
