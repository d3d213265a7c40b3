"""Parity of the HIP path (through the C ABI) against the CPU oracle, FP64 plain path.

Bar: integer results (supercell membership, per-cell counts, particle order) bit-exact;
FP64 fields, momenta and positions within 1e-12 relative (max |diff| / max |value|) after
100 steps; single-step quantities that do not depend on J-sum order bit-exact.
"""
import numpy as np
import pytest

import oracle as O
import pica_amd as P
from helpers import courant_dt, plasma, random_fields, random_particles, rel_err

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case] " + msg, flush=True)


def _gpu_grid_from(ctx, og, prec=P.F64, ghost=0):
    g = P.YeeGrid(ctx, og.s, og.origin, og.step, prec=prec, ghost=ghost)
    g.set_all(og.f)
    return g


@pytest.mark.parametrize("prec", [P.F64, P.F32])
def test_yee_update_matches_pica_semantics(ctx, prec):
    _say("yee prec=%d" % prec)
    rng = np.random.default_rng(11)
    dt_ = np.float64 if prec == P.F64 else np.float32
    og = O.Grid((10, 13, 9), (-2.0, -2.0, -2.0), (1.0, 0.5, 2.0), dtype=dt_)
    og.f[...] = rng.standard_normal(og.f.shape)
    g = _gpu_grid_from(ctx, og, prec)
    dt = 1.3e-11
    O.yee_update_b(og, dt * 0.5)
    P.YeeSolver().updateB(g, dt * 0.5)
    O.yee_update_e(og, dt)
    P.YeeSolver().updateE(g, dt)
    got = g.get_all()
    assert np.array_equal(got, og.f), "Yee update is not bit-identical to the oracle"


def test_ghost_fill_and_fold(ctx):
    _say("ghost fill/fold")
    rng = np.random.default_rng(12)
    og = O.Grid.periodic((6, 8, 5))
    og.f[...] = rng.standard_normal(og.f.shape)
    g = P.YeeGrid.periodic(ctx, (6, 8, 5))
    g.set_all(og.f)
    for c in (O.JX, O.JY, O.JZ):
        O.ghost_fold(og.f[c])
        O.ghost_fill(og.f[c])
    for c in range(6):
        O.ghost_fill(og.f[c])
    g.ghost_ops(P.FOLD_J | P.FILL_J | P.FILL_E | P.FILL_B)
    assert np.array_equal(g.get_all(), og.f)


def test_upload_buckets_like_pica(ctx):
    _say("upload bucketing")
    rng = np.random.default_rng(13)
    ncell, ncps = (8, 12, 4), (4, 4, 2)
    x, u = random_particles(rng, ncell, 3, uth=0.05)
    perm = rng.permutation(x.shape[0])
    x, u = x[perm], u[perm]
    w = rng.random(x.shape[0]) + 0.5
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps)
    ens.add(x, u, w)
    gx, gu, gw, goff = ens.particles()
    oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    assert np.array_equal(goff, oe.off)
    assert np.array_equal(gx, oe.positions()) and np.array_equal(gu, oe.momenta()) and np.array_equal(gw, oe.w)
    counts, cells = ens.cells()
    ocells = O.supercell_index((0, 0, 0), ncell, ncell, (1, 1, 1), oe.positions())
    lin = (ocells[:, 0] * ncell[1] + ocells[:, 1]) * ncell[2] + ocells[:, 2]
    assert np.array_equal(cells, lin)
    assert np.array_equal(counts.ravel(), np.bincount(lin, minlength=int(np.prod(ncell))))


def test_upload_rejects_outside_particles(ctx):
    _say("upload error path")
    ens = P.EnsembleSupercells(ctx, (4, 4, 4), (1.0, 1.0, 1.0), (2, 2, 2))
    x = np.array([[1.0, 1.0, 4.5]])
    with pytest.raises(P.PicaError) as ei:
        ens.add(x, np.zeros((1, 3)), np.ones(1))
    assert ei.value.code == P.PK_E_ARG


def _setup_pair(ctx, ncell, ncps, ppc, seed, shape, dep, uth=0.02, drift=0.0, fields=True):
    rng = np.random.default_rng(seed)
    dt = courant_dt()
    n0, w0 = plasma(dt, ppc)
    x, u = random_particles(rng, ncell, ppc, uth=uth, drift=drift)
    w = np.full(x.shape[0], w0)
    og = O.Grid.periodic(ncell)
    if fields:
        random_fields(rng, og, 3e-3, 3e-3)
    oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    g = P.YeeGrid.periodic(ctx, ncell)
    g.set_all(og.f)
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps)
    ens.add(x, u, w)
    return dt, og, oe, g, ens


@pytest.mark.parametrize("shape,dep", [(P.CIC, P.DIRECT), (P.CIC, P.ESIRKEPOV), (P.TSC, P.DIRECT),
                                       (P.TSC, P.ESIRKEPOV)])
def test_one_step_particles_bitexact(ctx, shape, dep):
    _say("one step shape=%d dep=%d" % (shape, dep))
    dt, og, oe, g, ens = _setup_pair(ctx, (8, 8, 12), (4, 4, 4), 6, 21 + 2 * shape + dep, shape, dep, uth=0.1)
    O.step(og, oe, dt, shape, dep)
    P.step(ens, g, dt, shape, dep)
    gx, gu, gw, goff = ens.particles()
    # particles depend only on the (identical) input fields: bit-exact, same order
    assert np.array_equal(goff, oe.off)
    assert np.array_equal(gx, oe.positions())
    assert np.array_equal(gu, oe.momenta())
    assert np.array_equal(gw, oe.w)
    # J and the updated fields differ only by the order of the J sums
    f = g.get_all()
    for c in range(9):
        assert rel_err(f[c], og.f[c]) < 1e-13, "component %d" % c


@pytest.mark.parametrize("shape,dep", [(P.CIC, P.DIRECT), (P.CIC, P.ESIRKEPOV), (P.TSC, P.DIRECT),
                                       (P.TSC, P.ESIRKEPOV)])
def test_twenty_steps_parity(ctx, shape, dep):
    _say("20 steps shape=%d dep=%d" % (shape, dep))
    dt, og, oe, g, ens = _setup_pair(ctx, (16, 8, 8), (4, 4, 4), 8, 31 + shape, shape, dep, uth=0.05, drift=0.1)
    for _ in range(20):
        O.step(og, oe, dt, shape, dep)
    P.run(ens, g, dt, 20, shape, dep)
    gx, gu, gw, goff = ens.particles()
    assert np.array_equal(goff, oe.off)
    assert rel_err(gx, oe.positions()) < 1e-12
    assert rel_err(gu, oe.momenta()) < 1e-12
    f = g.get_all()
    for c in range(9):
        assert rel_err(f[c], og.f[c]) < 1e-12, "component %d" % c


def test_thermal32_cic_direct_100_steps(ctx):
    """32^3 periodic thermal plasma, 8 ppc, CIC, direct deposition, 100 steps, FP64."""
    _say("thermal 32^3 x 8 ppc CIC/direct 100 steps")
    ncell, ncps, ppc = (32, 32, 32), (4, 4, 4), 8
    dt = courant_dt()
    n0, w0 = plasma(dt, ppc)
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps)
    ens.load_thermal(ppc, 0.01, 0.0, w0, 20240601)
    x, u, w, off = ens.particles()
    oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    assert np.array_equal(oe.off, off)
    og = O.Grid.periodic(ncell)
    g = P.YeeGrid.periodic(ctx, ncell)
    for _ in range(100):
        O.step(og, oe, dt, O.CIC, O.DIRECT)
    P.run(ens, g, dt, 100, P.CIC, P.DIRECT)
    gx, gu, gw, goff = ens.particles()
    assert np.array_equal(goff, oe.off), "supercell membership differs"
    counts, cells = ens.cells()
    ocells = O.supercell_index((0, 0, 0), ncell, ncell, (1, 1, 1), oe.positions())
    lin = (ocells[:, 0] * ncell[1] + ocells[:, 1]) * ncell[2] + ocells[:, 2]
    assert np.array_equal(cells, lin), "cell indices differ"
    assert np.array_equal(counts.ravel(), np.bincount(lin, minlength=int(np.prod(ncell))))
    ex, eu = rel_err(gx, oe.positions()), rel_err(gu, oe.momenta())
    f = g.get_all()
    ef = [rel_err(f[c], og.f[c]) for c in range(9)]
    print("thermal 32^3 CIC/direct rel errors: x %.3g u %.3g fields %s" % (ex, eu, " ".join("%.3g" % v for v in ef)), flush=True)
    assert ex < 1e-12 and eu < 1e-12
    assert max(ef) < 1e-12


def test_two_stream_32cube_tsc_esirkepov_100_steps(ctx):
    """The headline plasma's physics on a 32^3 grid: 16 ppc, Maxwellian 0.01 plus a two-stream +-0.1c drift
    in x on alternate particles, TSC, Esirkepov, 4^3 supercells re-sorted every step, 100 FP64
    steps.  Membership, cell indices and per-cell counts bit-exact; x, u, E, B, J within 1e-12."""
    _say("two-stream 32^3 x 16 ppc TSC/Esirkepov 100 steps")
    ncell, ncps, ppc = (32, 32, 32), (4, 4, 4), 16
    dt = courant_dt()
    n0, w0 = plasma(dt, ppc)
    drift = 0.1 / np.sqrt(1.0 - 0.1 * 0.1)
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps)
    ens.load_thermal(ppc, 0.01, drift, w0, 20240602)
    x, u, w, off = ens.particles()
    oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    assert np.array_equal(oe.off, off)
    og = O.Grid.periodic(ncell)
    g = P.YeeGrid.periodic(ctx, ncell)
    for _ in range(100):
        O.step(og, oe, dt, O.TSC, O.ESIRKEPOV)
    P.run(ens, g, dt, 100, P.TSC, P.ESIRKEPOV)
    gx, gu, gw, goff = ens.particles()
    assert np.array_equal(goff, oe.off), "supercell membership differs"
    counts, cells = ens.cells()
    ocells = O.supercell_index((0, 0, 0), ncell, ncell, (1, 1, 1), oe.positions())
    lin = (ocells[:, 0] * ncell[1] + ocells[:, 1]) * ncell[2] + ocells[:, 2]
    assert np.array_equal(cells, lin), "cell indices differ"
    assert np.array_equal(counts.ravel(), np.bincount(lin, minlength=int(np.prod(ncell))))
    assert np.array_equal(gw, oe.w), "particle order differs"
    ex, eu = rel_err(gx, oe.positions()), rel_err(gu, oe.momenta())
    f = g.get_all()
    ef = [rel_err(f[c], og.f[c]) for c in range(9)]
    print("two-stream 32^3 TSC/Esirkepov rel errors: x %.3g u %.3g fields %s" % (ex, eu, " ".join("%.3g" % v for v in ef)),
          flush=True)
    assert ex < 1e-12 and eu < 1e-12
    assert max(ef) < 1e-12


def test_nonfinite_particle_reports_error(ctx):
    _say("non-finite particle error path")
    ncell = (4, 4, 4)
    g = P.YeeGrid.periodic(ctx, ncell)
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), (2, 2, 2))
    x = np.array([[0.5, 0.5, 0.5], [1.5, 2.5, 3.5]])
    u = np.array([[0.0, 0.0, 0.0], [np.nan, 0.0, 0.0]])
    ens.add(x, u, np.ones(2))
    with pytest.raises(P.PicaError) as ei:
        P.step(ens, g, courant_dt(), P.CIC, P.DIRECT)
    assert ei.value.code == P.PK_E_STATE
