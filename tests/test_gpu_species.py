"""Several species on one grid (pica's ParticleTypes table, Particle.h:26-31, BorisPusher.h:26-29):
electrons and mobile ions, one ensemble each, stepped by pk_step_species.

Checks: parity with the oracle running the same two-species step (each ensemble's particle stage
into one J, then the field stage), and Gauss's law with the summed charge density of both
species at round-off for the charge-conserving depositor."""
import math

import numpy as np
import pytest

import oracle as O
import pica_amd as P
from helpers import courant_dt, plasma, random_particles, rel_err

pytestmark = pytest.mark.gpu

MASS_RATIO = 100.0        # light ions, so that they move within a few steps


def _say(msg):
    print("[case] " + msg, flush=True)


def _species(ctx):
    me, qe = O.E_MASS, O.E_CHARGE
    ctx.set_species([me, MASS_RATIO * me], [qe, -qe])
    return (me, qe), (MASS_RATIO * me, -qe)


def _two_plasmas(ctx, ncell, ppc, seed, prec=P.F64):
    rng = np.random.default_rng(seed)
    dt = courant_dt()
    _, w0 = plasma(dt, ppc)
    xe, ue = random_particles(rng, ncell, ppc, uth=0.05, drift=0.1)
    xi, ui = random_particles(rng, ncell, ppc, uth=0.05 / math.sqrt(MASS_RATIO), drift=0.0)
    ui[:, 1] += 0.02                                  # an ion beam across y
    we, wi = np.full(xe.shape[0], w0), np.full(xi.shape[0], w0)
    return dt, (xe, ue, we), (xi, ui, wi)


@pytest.mark.parametrize("shape,dep", [(P.CIC, P.DIRECT), (P.TSC, P.ESIRKEPOV)])
def test_two_species_match_oracle(ctx, shape, dep):
    _say("two species shape=%d dep=%d" % (shape, dep))
    (me, qe), (mi, qi) = _species(ctx)
    try:
        ncell, ncps = (16, 8, 8), (4, 4, 4)
        dt, ele, ion = _two_plasmas(ctx, ncell, 4, 71 + shape)
        og = O.Grid.periodic(ncell)
        oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), *ele, q=qe, m=me)
        oi = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), *ion, q=qi, m=mi)
        g = P.YeeGrid.periodic(ctx, ncell)
        e = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps, type_index=0)
        i = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps, type_index=1)
        e.add(*ele)
        i.add(*ion)
        for _ in range(10):
            og.f[O.JX:O.JZ + 1] = 0.0
            O.particle_stage(og, oe, dt, shape, dep)
            O.particle_stage(og, oi, dt, shape, dep)
            O.field_stage(og, dt)
            O.resort(oe)
            O.resort(oi)
            P.step_species([e, i], g, dt, shape, dep)
        errs = {}
        for name, ge, oo in (("electrons", e, oe), ("ions", i, oi)):
            x, u, w, off = ge.particles()
            assert np.array_equal(off, oo.off), name
            errs[name] = (rel_err(x, oo.positions()), rel_err(u, oo.momenta()))
        f = g.get_all()
        ef = max(rel_err(f[c], og.f[c]) for c in range(9))
        print("two species: %s fields %.3g" % (errs, ef), flush=True)
        assert max(max(v) for v in errs.values()) < 1e-12 and ef < 1e-12
    finally:
        ctx.set_species([O.E_MASS], [O.E_CHARGE])


def test_two_species_gauss_law(ctx):
    _say("two species gauss")
    _species(ctx)
    try:
        ncell, ncps = (16, 16, 8), (4, 4, 4)
        dt, ele, ion = _two_plasmas(ctx, ncell, 4, 81)
        g = P.YeeGrid.periodic(ctx, ncell)
        e = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps, type_index=0)
        i = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps, type_index=1)
        e.add(*ele)
        i.add(*ion)
        G = O.GHOST

        def residual():
            rho = 0.0
            for ens in (e, i):
                P.deposit_rho(ens, g, P.TSC)
                rho = rho + g.get(P.RHO)
            ex, ey, ez = (g.get(c) for c in (P.EX, P.EY, P.EZ))
            div = O.physical_nodes(O.divergence(ex, ey, ez, (1.0, 1.0, 1.0)))
            r = div - 4.0 * O.PI * rho[G:-G, G:-G, G:-G]
            return r, np.abs(4.0 * O.PI * rho[G:-G, G:-G, G:-G]).max()

        r0, _ = residual()
        for _ in range(20):
            P.step_species([e, i], g, dt, P.TSC, P.ESIRKEPOV)
        r1, scale = residual()
        drift = np.abs(r1 - r0).max() / scale
        print("two species gauss drift %.3g" % drift, flush=True)
        assert drift < 1e-10
    finally:
        ctx.set_species([O.E_MASS], [O.E_CHARGE])
