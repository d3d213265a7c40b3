"""Batched form: many (ensemble, grid) pairs of uneven sizes and particle counts stepped together
(one launch per stage, a captured HIP graph per step) give the results of stepping each pair on
its own with pk_run.  Particle order inside a supercell is deterministic, so particles are
compared slot by slot; the only differences are the summation order of global J atomics (FP64:
round-off; FP32: FP32 round-off), bounded per quantity below."""
import numpy as np
import pytest

import pica_amd as P
from helpers import courant_dt, plasma, random_particles, rel_err

pytestmark = pytest.mark.gpu

# (ncell, ppc, seed): uneven grids and particle counts, 4^3 supercells
SPECS = [((8, 8, 12), 8, 11), ((12, 8, 8), 4, 12), ((8, 12, 16), 12, 13), ((4, 8, 4), 6, 14)]
TOL = {P.F64: 1e-12, P.F32: 2e-5}


def _say(msg):
    print("[case] " + msg, flush=True)


@pytest.fixture(autouse=True)
def _no_interleave(ctx):
    """these tests compare batched with one-by-one stepping without the periodic interleave (its
    own test below turns it on for both)"""
    period = ctx.get_option(P.OPT_INTERLEAVE)
    ctx.set_option(P.OPT_INTERLEAVE, 0)
    yield
    ctx.set_option(P.OPT_INTERLEAVE, period)


def _pair(ctx, prec, spec, uth=0.05, slack=None):
    ncell, ppc, seed = spec
    rng = np.random.default_rng(seed)
    dt = courant_dt()
    _, w0 = plasma(dt, ppc)
    x, u = random_particles(rng, ncell, ppc, uth=uth, drift=0.1)
    w = (w0 * (1.0 + np.arange(x.shape[0]) * 2.0 ** -20)).astype(np.float32).astype(np.float64)
    g = P.YeeGrid.periodic(ctx, ncell, prec=prec)
    e = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), (4, 4, 4), prec=prec)
    if slack is not None:
        e.set_slack(*slack)
    e.add(x, u, w)
    return e, g


def _pairs(ctx, prec, specs=SPECS, **kw):
    return [_pair(ctx, prec, s, **kw) for s in specs]


def _state(e, g):
    x, u, w, off = e.particles()
    return x, u, w, off, g.get_all()


def _compare(prec, a, b, what):
    xa, ua, wa, oa, fa = a
    xb, ub, wb, ob, fb = b
    assert np.array_equal(oa, ob), "%s: supercell offsets differ" % what
    assert np.array_equal(wa, wb), "%s: particle order differs" % what
    errs = {"x": rel_err(xa, xb), "u": rel_err(ua, ub),
            "E": max(rel_err(fa[c], fb[c]) for c in (P.EX, P.EY, P.EZ)),
            "B": max(rel_err(fa[c], fb[c]) for c in (P.BX, P.BY, P.BZ)),
            "J": max(rel_err(fa[c], fb[c]) for c in (P.JX, P.JY, P.JZ))}
    for k, v in errs.items():
        assert v < TOL[prec], "%s: %s differs by %.3g" % (what, k, v)
    return errs


def _check(ctx, prec, shape, dep, nsteps, knobs=None, specs=SPECS, **kw):
    dt = courant_dt()
    mine = _pairs(ctx, prec, specs, **kw)
    ref = _pairs(ctx, prec, specs, **kw)
    b = P.Batch(ctx, mine)
    for k, v in (knobs or {}).items():
        b.set(k, v)
    b.run(dt, nsteps, shape, dep)
    for e, g in ref:
        P.run(e, g, dt, nsteps, shape, dep)
    worst = {}
    for i, ((e, g), (re_, rg)) in enumerate(zip(mine, ref)):
        errs = _compare(prec, _state(e, g), _state(re_, rg), "ensemble %d" % i)
        worst = {k: max(v, worst.get(k, 0.0)) for k, v in errs.items()}
    print("batch vs one-by-one (prec=%d shape=%d dep=%d knobs=%s): %s" % (prec, shape, dep, knobs, worst), flush=True)
    b.close()


@pytest.mark.parametrize("prec,shape,dep", [(P.F64, P.TSC, P.ESIRKEPOV), (P.F64, P.CIC, P.DIRECT),
                                            (P.F32, P.TSC, P.ESIRKEPOV), (P.F32, P.CIC, P.ESIRKEPOV)])
def test_batch_matches_one_by_one(ctx, prec, shape, dep):
    _say("batch prec=%d shape=%d dep=%d" % (prec, shape, dep))
    _check(ctx, prec, shape, dep, 6)


@pytest.mark.parametrize("wg,order,graph", [(0, 0, 0), (0, 1, 1), (1, 0, 1), (2, 1, 0)])
def test_batch_knobs(ctx, wg, order, graph):
    """persistent or one-workgroup-per-item launches, both item orders, graph or eager"""
    _say("batch knobs wg=%d order=%d graph=%d" % (wg, order, graph))
    _check(ctx, P.F32, P.TSC, P.ESIRKEPOV, 5, {P.BATCH_WG_PER_CU: wg, P.BATCH_ORDER: order, P.BATCH_GRAPH: graph})


def test_batch_graph_replays(ctx):
    """the captured step replayed one step at a time, with eager work of the program (downloads,
    a diagnostic, a single-ensemble resort) and synchronisations between the replays; every
    replay equals the one-by-one step"""
    _say("batch graph replays")
    dt = courant_dt()
    mine = _pairs(ctx, P.F64)
    ref = _pairs(ctx, P.F64)
    b = P.Batch(ctx, mine)
    b.set(P.BATCH_GRAPH, 1)
    for k in range(4):
        b.run(dt, 1, P.TSC, P.ESIRKEPOV)
        for e, g in ref:
            P.step(e, g, dt, P.TSC, P.ESIRKEPOV)
        ctx.sync()
        for i, ((e, g), (re_, rg)) in enumerate(zip(mine, ref)):
            _compare(P.F64, _state(e, g), _state(re_, rg), "replay %d ensemble %d" % (k, i))
        P.diag(mine[0][0], mine[0][1], P.TSC)
        P.diag(ref[0][0], ref[0][1], P.TSC)
    b.close()


def test_batch_slack_overflow_recovery(ctx):
    """no slack and a hot plasma: inserts overflow in most steps; the batch stops its queued steps,
    rebuilds the overflowed ensembles and resumes, ending where the one-by-one run (which
    rebuilds after each overflowing insert) ends"""
    _say("batch overflow recovery")
    ctx.timer_reset()
    ctx.timer_enable(True)
    try:
        _check(ctx, P.F64, P.TSC, P.ESIRKEPOV, 6, uth=0.2, slack=(0.0, 0))
        _, rebuilds = ctx.timer("rebuild")
    finally:
        ctx.timer_enable(False)
    print("rebuilds during the overflow test: %d" % rebuilds, flush=True)
    assert rebuilds > 0


def test_batch_autotune_restores_state(ctx):
    """the autotuner's trials leave the state as it was; the chosen setting then runs correctly"""
    _say("batch autotune")
    dt = courant_dt()
    mine = _pairs(ctx, P.F32)
    ref = _pairs(ctx, P.F32)
    b = P.Batch(ctx, mine)
    before = [_state(e, g) for e, g in mine]
    ms = b.autotune(dt, P.TSC, P.ESIRKEPOV, trial_steps=2)
    after = [_state(e, g) for e, g in mine]
    for s0, s1 in zip(before, after):
        for a0, a1 in zip(s0, s1):
            assert np.array_equal(a0, a1)
    knobs = [b.get(k) for k in (P.BATCH_WG_PER_CU, P.BATCH_ORDER, P.BATCH_GRAPH)]
    print("autotune: %.4f ms/step with knobs %s" % (ms, knobs), flush=True)
    assert ms > 0 and knobs[0] in (0, 1, 2, 4, 8) and knobs[1] in (0, 1) and knobs[2] in (0, 1)
    b.run(dt, 3, P.TSC, P.ESIRKEPOV)
    for e, g in ref:
        P.run(e, g, dt, 3, P.TSC, P.ESIRKEPOV)
    for i, ((e, g), (re_, rg)) in enumerate(zip(mine, ref)):
        _compare(P.F32, _state(e, g), _state(re_, rg), "ensemble %d" % i)
    b.close()


# a mix of supercell sizes: 40 ppc puts ~2560 particles in a 4^3 supercell (the interleave's
# form for 2049-4096 particles runs beside the one for up to 2048)
MIXED = [((8, 8, 8), 40, 21), ((8, 8, 12), 8, 11)]


@pytest.mark.parametrize("period,graph,uth,slack,specs", [(2, 1, 0.05, None, SPECS), (3, 0, 0.05, None, SPECS),
                                                          (1, 1, 0.2, (0.0, 0), SPECS), (2, 1, 0.05, None, MIXED)])
def test_batch_interleave_matches_one_by_one(ctx, period, graph, uth, slack, specs):
    """with the periodic interleave on, the batch reorders each FP32 ensemble after the same
    re-sorts as one-by-one stepping -- counted per ensemble (one ensemble starts a step ahead),
    and also after an overflow rebuild (no slack, hot plasma), and with supercells of both size
    classes -- and ends in the same state, particle order included"""
    _say("batch interleave period=%d graph=%d uth=%g nspecs=%d" % (period, graph, uth, len(specs)))
    ctx.set_option(P.OPT_INTERLEAVE, period)
    dt = courant_dt()
    mine = _pairs(ctx, P.F32, specs, uth=uth, slack=slack)
    ref = _pairs(ctx, P.F32, specs, uth=uth, slack=slack)
    for e, g in (mine[0], ref[0]):
        P.step(e, g, dt, P.TSC, P.ESIRKEPOV)
    b = P.Batch(ctx, mine)
    b.set(P.BATCH_GRAPH, graph)
    ctx.timer_reset()
    ctx.timer_enable(True)
    try:
        b.run(dt, 7, P.TSC, P.ESIRKEPOV)
        for e, g in ref:
            P.run(e, g, dt, 7, P.TSC, P.ESIRKEPOV)
        ctx.sync()
        _, rebuilds = ctx.timer("rebuild")
    finally:
        ctx.timer_enable(False)
    worst = {}
    for i, ((e, g), (re_, rg)) in enumerate(zip(mine, ref)):
        errs = _compare(P.F32, _state(e, g), _state(re_, rg), "ensemble %d" % i)
        worst = {k: max(v, worst.get(k, 0.0)) for k, v in errs.items()}
    print("batch vs one-by-one with interleave period %d: %s, rebuilds %d" % (period, worst, rebuilds), flush=True)
    if slack is not None:
        assert rebuilds > 0
    b.close()
