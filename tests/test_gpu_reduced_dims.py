"""1D and 2D grids: a 3D grid whose reduced dimensions have one cell (pica's YeeGrid<One / Two>
x and x-y layouts, YeeSolver.h:41-128, FieldInterpolatorYeeGrid.h:51-133,
CurrentDepositorYeeGrid.h:52-122), all of whose periodic ghosts image that cell, so every
derivative along it vanishes.

Checks:
  - parity with the oracle (the same 3D code with one-cell edges): particles bit-exact after one
    step, fields and particles within 1e-12 after 20 steps, FP64;
  - the 2D run equals a 3D run of the same plasma replicated in 4 z layers (fields z-independent
    and equal to the 2D fields, x-y motion identical): pins what the one-cell edge means;
  - the FP32 split form against the FP64 oracle in 2D;
  - the denoiser and the projection on a 2D grid keep Gauss's law.
"""
import numpy as np
import pytest

import oracle as O
import pica_amd as P
from helpers import courant_dt, plasma, random_particles, rel_err

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case] " + msg, flush=True)


def _pair(ctx, ncell, ncps, ppc, seed, prec=P.F64, uth=0.05, drift=0.1):
    rng = np.random.default_rng(seed)
    dt = courant_dt()
    _, w0 = plasma(dt, ppc)
    x, u = random_particles(rng, ncell, ppc, uth=uth, drift=drift)
    u[:, 2] += 0.05 * np.sin(2 * np.pi * x[:, 0] / ncell[0])   # motion across the one-cell edge too
    w = np.full(x.shape[0], w0)
    og = O.Grid.periodic(ncell)
    oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    g = P.YeeGrid.periodic(ctx, ncell, prec=prec)
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps, prec=prec)
    ens.add(x, u, w)
    return dt, og, oe, g, ens, (x, u, w)


CASES = [((16, 16, 1), (4, 4, 1)), ((32, 1, 1), (4, 1, 1)), ((8, 1, 16), (4, 1, 4))]


@pytest.mark.parametrize("ncell,ncps", CASES)
@pytest.mark.parametrize("shape,dep", [(P.CIC, P.DIRECT), (P.TSC, P.ESIRKEPOV)])
def test_reduced_dims_match_oracle(ctx, ncell, ncps, shape, dep):
    _say("reduced dims %s shape=%d dep=%d" % (ncell, shape, dep))
    dt, og, oe, g, ens, _ = _pair(ctx, ncell, ncps, 8, 41 + shape, uth=0.05)
    O.step(og, oe, dt, shape, dep)
    P.step(ens, g, dt, shape, dep)
    gx, gu, gw, goff = ens.particles()
    assert np.array_equal(goff, oe.off)
    assert np.array_equal(gx, oe.positions()) and np.array_equal(gu, oe.momenta())
    for _ in range(19):
        O.step(og, oe, dt, shape, dep)
    P.run(ens, g, dt, 19, shape, dep)
    gx, gu, gw, goff = ens.particles()
    assert np.array_equal(goff, oe.off)
    ex, eu = rel_err(gx, oe.positions()), rel_err(gu, oe.momenta())
    f = g.get_all()
    ef = max(rel_err(f[c], og.f[c]) for c in range(9))
    print("reduced dims %s: x %.3g u %.3g fields %.3g" % (ncell, ex, eu, ef), flush=True)
    assert ex < 1e-12 and eu < 1e-12 and ef < 1e-12


@pytest.mark.parametrize("shape,dep", [(P.CIC, P.DIRECT), (P.TSC, P.ESIRKEPOV)])
def test_2d_equals_replicated_3d(ctx, shape, dep):
    _say("2D vs replicated 3D shape=%d dep=%d" % (shape, dep))
    nz = 4
    dt, _, _, g2, e2, (x, u, w) = _pair(ctx, (8, 8, 1), (4, 4, 1), 8, 51 + shape)
    n = x.shape[0]
    x3 = np.concatenate([x + np.array([0.0, 0.0, float(k)]) for k in range(nz)])
    u3 = np.tile(u, (nz, 1))
    w3 = np.tile(w, nz)
    g3 = P.YeeGrid.periodic(ctx, (8, 8, nz))
    e3 = P.EnsembleSupercells(ctx, (8, 8, nz), (1.0, 1.0, 1.0), (4, 4, 4))
    e3.add(x3, u3, w3)
    P.run(e2, g2, dt, 10, shape, dep)
    P.run(e3, g3, dt, 10, shape, dep)
    G = O.GHOST
    worst = 0.0
    for c in range(9):
        a2 = g2.get(c)[G:-G, G:-G, G:-G]
        a3 = g3.get(c)[G:-G, G:-G, G:-G]
        for k in range(nz):
            worst = max(worst, rel_err(a3[:, :, k:k + 1], a2) if np.abs(a2).max() > 0 else float(np.abs(a3).max()))
    x2, v2, _, _ = e2.particles()
    x3g, v3, _, _ = e3.particles()
    # match particles by (x, y, z mod 1, u): the 3D copies of one 2D particle move identically
    k2 = np.lexsort((x2[:, 2] % 1.0, x2[:, 1], x2[:, 0]))
    r3 = np.stack([x3g[:, 0], x3g[:, 1], x3g[:, 2] % 1.0], 1)
    k3 = np.lexsort((r3[:, 2], r3[:, 1], r3[:, 0]))
    xy_err = rel_err(np.repeat(np.stack([x2[k2, 0], x2[k2, 1], x2[k2, 2] % 1.0], 1), nz, 0), r3[k3])
    print("2D vs 3D replicated: fields %.3g positions %.3g" % (worst, xy_err), flush=True)
    assert x3g.shape[0] == nz * n
    assert worst < 1e-12 and xy_err < 1e-12
    for o in (g3, e3):
        o.close()


def test_split_2d_matches_fp64_oracle(ctx):
    _say("FP32 split 2D")
    from test_gpu_split import _compare
    _compare(ctx, P.TSC, P.ESIRKEPOV, (16, 16, 1), (4, 4, 1), 8, 10, 61)


@pytest.mark.parametrize("prec,tol", [(P.F64, 1e-10), (P.F32, 2e-4)])
def test_denoised_step_2d_keeps_gauss_law(ctx, prec, tol):
    _say("denoised step 2D prec=%d" % prec)
    ncell, dt = (16, 16, 1), courant_dt()
    n0, w0 = plasma(dt, 4)
    g = P.YeeGrid.periodic(ctx, ncell, prec=prec)
    e = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), (4, 4, 1), prec=prec)
    e.load_modes(4, 0.02, 0.0, w0, 13, [[1, 0, 0, 0.05, 0, 0.02, 0.3], [0, 1, 0, 0.01, 0.03, 0, 1.1]])
    dn = P.Denoiser(ctx)
    dn.init(5)
    dn.set_scale(1.0 / (abs(P.ELECTRON_CHARGE) * n0 * P.C * 0.05))
    P.diag_set_reference(e, g, P.TSC)
    for _ in range(5):
        P.step_denoised(e, g, dn, dt, P.TSC, P.ESIRKEPOV)
    d = P.diag(e, g, P.TSC)
    drift = d[P.DIAG_GAUSS_DRIFT] / d[P.DIAG_RHO_MAX]
    print("denoised 2D prec=%d: gauss drift %.3g" % (prec, drift), flush=True)
    assert drift < tol
    dn.close()


def test_projection_2d_matches_oracle(ctx):
    """the continuity projection on a 2D grid (no FFT pass along the one-cell edge) equals the
    numpy FFT solve and removes the continuity residual"""
    from oracle import denoiser as D
    _say("projection 2D")
    ncell, dt = (16, 8, 1), courant_dt()
    rng = np.random.default_rng(7)
    _, w0 = plasma(dt, 4)
    x, u = random_particles(rng, ncell, 4, uth=0.05, drift=0.1)
    g = P.YeeGrid.periodic(ctx, ncell)
    e = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), (4, 4, 1))
    e.add(x, u, np.full(x.shape[0], w0))
    P.run(e, g, dt, 2, P.TSC, P.DIRECT)
    P.deposit_rho(e, g, P.TSC)
    rho0 = g.get(P.RHO).astype(np.float64)
    P.rho_save(e, g, P.TSC)
    for c in (P.JX, P.JY, P.JZ):
        g.zero(c)
    P.particle_step(e, g, dt, P.TSC, P.DIRECT)
    g.ghost_ops(P.FOLD_J | P.FILL_J)
    j_raw = [g.get(c).astype(np.float64) for c in (P.JX, P.JY, P.JZ)]
    P.project_continuity(e, g, dt, P.TSC)
    j_new = [g.get(c).astype(np.float64) for c in (P.JX, P.JY, P.JZ)]
    P.deposit_rho(e, g, P.TSC)
    rho1 = g.get(P.RHO).astype(np.float64)
    G = O.GHOST
    phys = lambda a: a[G:-G, G:-G, G:-G]
    ref = D.project_continuity(*[phys(a) for a in j_raw], phys(rho0), phys(rho1), dt, (1.0, 1.0, 1.0))
    err = max(rel_err(phys(a), b) for a, b in zip(j_new, ref))
    scale = np.abs(D.divergence_periodic(*[phys(a) for a in j_raw], (1, 1, 1))).max()
    after = D.divergence_periodic(*[phys(a) for a in j_new], (1, 1, 1)) + (phys(rho1) - phys(rho0)) / dt
    print("projection 2D: vs oracle %.3g, residual %.3g" % (err, np.abs(after).max() / scale), flush=True)
    assert err < 1e-12 and np.abs(after).max() / scale < 1e-11
