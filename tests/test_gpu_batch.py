"""Batched form: many (ensemble, grid) pairs of uneven sizes and particle counts stepped together
(one launch per stage, a captured HIP graph per step) give the results of stepping each pair on
its own with pk_run.  Particle order inside a supercell is deterministic, so particles are
compared slot by slot; the only differences are the summation order of global J atomics (FP64:
round-off; FP32: FP32 round-off), bounded per quantity below."""
import numpy as np
import pytest

import oracle as O
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


@pytest.mark.parametrize("wg,order,graph,block", [(0, 0, 0, 256), (0, 1, 1, 256), (1, 0, 1, 256), (2, 1, 0, 256),
                                                  (0, 1, 1, 128), (2, 0, 1, 128)])
def test_batch_knobs(ctx, wg, order, graph, block):
    """persistent or one-workgroup-per-item launches, both item orders, graph or eager, workgroups of
    four or two wavefronts"""
    _say("batch knobs wg=%d order=%d graph=%d block=%d" % (wg, order, graph, block))
    _check(ctx, P.F32, P.TSC, P.ESIRKEPOV, 5, {P.BATCH_WG_PER_CU: wg, P.BATCH_ORDER: order, P.BATCH_GRAPH: graph,
                                               P.BATCH_BLOCK: block})


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


@pytest.mark.parametrize("prec", [P.F64, P.F32])
def test_batch_autotune_restores_state(ctx, prec):
    """the autotuner's trials (supercell sizes, interleave periods, launch settings) leave every
    particle and field as they were, re-bucketed to the chosen supercell size exactly as
    pk_ens_set_supercells does; the chosen setting then runs like one-by-one stepping with the
    same supercells and interleave period"""
    _say("batch autotune prec=%d" % prec)
    dt = courant_dt()
    mine = _pairs(ctx, prec)
    ref = _pairs(ctx, prec)
    b = P.Batch(ctx, mine)
    ms = b.autotune(dt, P.TSC, P.ESIRKEPOV, trial_steps=2)
    knobs = {k: b.get(k) for k in (P.BATCH_WG_PER_CU, P.BATCH_ORDER, P.BATCH_GRAPH, P.BATCH_SUPERCELL,
                                   P.BATCH_INTERLEAVE, P.BATCH_BLOCK)}
    print("autotune: %.4f ms/step with knobs %s" % (ms, knobs), flush=True)
    assert ms > 0 and knobs[P.BATCH_WG_PER_CU] in (0, 1, 2, 4, 8) and knobs[P.BATCH_ORDER] in (0, 1)
    assert knobs[P.BATCH_GRAPH] in (0, 1) and knobs[P.BATCH_SUPERCELL] in (2, 4)
    assert knobs[P.BATCH_INTERLEAVE] in ((-1,) if prec == P.F64 else (0, 2, 3, 4, 6, 8))
    assert knobs[P.BATCH_BLOCK] in ((256,) if prec == P.F64 else (128, 256))
    sc = knobs[P.BATCH_SUPERCELL]
    for e, g in ref:
        e.set_supercells((sc, sc, sc))
    for i, ((e, g), (re_, rg)) in enumerate(zip(mine, ref)):
        for a0, a1 in zip(_state(e, g), _state(re_, rg)):
            assert np.array_equal(a0, a1), "ensemble %d changed by the autotuner" % i
    period = ctx.get_option(P.OPT_INTERLEAVE)
    if knobs[P.BATCH_INTERLEAVE] >= 0:
        ctx.set_option(P.OPT_INTERLEAVE, knobs[P.BATCH_INTERLEAVE])
    try:
        b.run(dt, 3, P.TSC, P.ESIRKEPOV)
        for e, g in ref:
            P.run(e, g, dt, 3, P.TSC, P.ESIRKEPOV)
    finally:
        ctx.set_option(P.OPT_INTERLEAVE, period)
    for i, ((e, g), (re_, rg)) in enumerate(zip(mine, ref)):
        _compare(prec, _state(e, g), _state(re_, rg), "ensemble %d" % i)
    b.close()


@pytest.mark.parametrize("prec", [P.F64, P.F32])
def test_batch_per_instance_form_factor(ctx, prec):
    """CIC and TSC instances in one batch (PK_SHAPE_PER_INSTANCE: one particle launch per form
    factor) give each instance's one-by-one result with its own form factor"""
    _say("batch per-instance shapes prec=%d" % prec)
    dt = courant_dt()
    shapes = [P.CIC, P.TSC, P.TSC, P.CIC]
    mine = _pairs(ctx, prec)
    ref = _pairs(ctx, prec)
    b = P.Batch(ctx, mine)
    for i, s in enumerate(shapes):
        b.set_shape(i, s)
    b.run(dt, 5, P.SHAPE_PER_INSTANCE, P.ESIRKEPOV)
    for (e, g), s in zip(ref, shapes):
        P.run(e, g, dt, 5, s, P.ESIRKEPOV)
    for i, ((e, g), (re_, rg)) in enumerate(zip(mine, ref)):
        _compare(prec, _state(e, g), _state(re_, rg), "ensemble %d (shape %d)" % (i, shapes[i]))
    # an unset shape is refused
    b2 = P.Batch(ctx, _pairs(ctx, prec, SPECS[:2]))
    b2.set_shape(0, P.TSC)
    with pytest.raises(P.PicaError) as ei:
        b2.run(dt, 1, P.SHAPE_PER_INSTANCE, P.ESIRKEPOV)
    assert ei.value.code == P.PK_E_ARG
    b2.close()
    b.close()


@pytest.mark.parametrize("ncps", [(2, 2, 2), (8, 8, 8), (2, 4, 8)])
def test_set_supercells_matches_the_oracle(ctx, ncps):
    """re-bucketing an FP64 ensemble into other supercells gives pica's bucketing of the same
    particles (membership and order bit-exact: stable by new supercell in the old order), and 6
    TSC/Esirkepov steps then match the oracle with those supercells at 1e-12"""
    _say("set_supercells ncps=%s" % (ncps,))
    ncell = (16, 16, 8)
    dt = courant_dt()
    _, w0 = plasma(dt, 8)
    e = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), (4, 4, 4))
    e.load_thermal(8, 0.02, 0.1, w0, 77)
    g = P.YeeGrid.periodic(ctx, ncell)
    P.run(e, g, dt, 2, P.TSC, P.ESIRKEPOV)
    x, u, w, off = e.particles()
    og = O.Grid.periodic(ncell)
    og.f[...] = g.get_all()[:9]
    e.set_supercells(ncps)
    oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    gx, gu, gw, goff = e.particles()
    assert np.array_equal(goff, oe.off) and np.array_equal(gx, oe.positions()) and np.array_equal(gu, oe.momenta())
    for _ in range(6):
        O.step(og, oe, dt, O.TSC, O.ESIRKEPOV)
    P.run(e, g, dt, 6, P.TSC, P.ESIRKEPOV)
    gx, gu, gw, goff = e.particles()
    assert np.array_equal(goff, oe.off)
    assert rel_err(gx, oe.positions()) < 1e-12 and rel_err(gu, oe.momenta()) < 1e-12
    f = g.get_all()
    assert max(rel_err(f[c], og.f[c]) for c in range(9)) < 1e-12
    e.close()
    g.close()


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


@pytest.mark.parametrize("prec,shape,dep,period,slack", [(P.F64, P.TSC, P.ESIRKEPOV, 0, None),
                                                         (P.F32, P.TSC, P.ESIRKEPOV, 3, None),
                                                         (P.F64, P.CIC, P.DIRECT, 0, (0.0, 2)),
                                                         (P.F32, P.CIC, P.ESIRKEPOV, 2, (0.0, 2))])
def test_run_captured_matches_run(ctx, prec, shape, dep, period, slack):
    """pk_run_captured (the step captured once, no host synchronisation between steps) gives
    pk_run's particles, order and fields, with the interleave at the same re-sort counts and with
    slack overflows rebuilt after the replay stops on the device"""
    _say("run captured prec=%d shape=%d dep=%d period=%d slack=%s" % (prec, shape, dep, period, slack))
    ctx.set_option(P.OPT_INTERLEAVE, period)
    dt = courant_dt()
    for spec in SPECS[:2]:
        (e, g), (re_, rg) = _pair(ctx, prec, spec, slack=slack), _pair(ctx, prec, spec, slack=slack)
        P.run_captured(e, g, dt, 9, shape, dep)
        P.run(re_, rg, dt, 9, shape, dep)
        _compare(prec, _state(e, g), _state(re_, rg), "ensemble %s" % (spec,))
        P.run_captured(e, g, dt, 0, shape, dep)              # no steps: nothing changes
        _compare(prec, _state(e, g), _state(re_, rg), "ensemble %s after 0 steps" % (spec,))


def test_run_captured_timing_32cube(ctx):
    """a thermal 32^3 x 8 ppc plasma (CIC, direct, FP64), 100 steps each way from the same
    particles: the captured run matches pk_run and is timed beside it"""
    _say("run captured timing, 32^3 x 8 ppc")
    import time
    dt = courant_dt()
    spec = ((32, 32, 32), 8, 21)
    runs = []
    for fn in (P.run, P.run_captured):
        e, g = _pair(ctx, P.F64, spec, uth=0.01)
        fn(e, g, dt, 2, P.CIC, P.DIRECT)
        ctx.sync()
        t0 = time.perf_counter()
        fn(e, g, dt, 100, P.CIC, P.DIRECT)
        ctx.sync()
        runs.append((time.perf_counter() - t0, _state(e, g)))
    print("32^3 x 8 ppc, 100 steps: pk_run %.2f ms per step, pk_run_captured %.2f ms per step"
          % (10.0 * runs[0][0], 10.0 * runs[1][0]), flush=True)
    _compare(P.F64, runs[1][1], runs[0][1], "32^3 x 8 ppc")
