"""Charge conservation of the Esirkepov depositor on the device, FP64 and FP32 split forms:
the discrete continuity residual (rho^{n+1} - rho^n)/dt + div J and the drift of the Gauss-law
residual div E - 4 pi rho stay at round-off; a direct depositor does not conserve (control).
Sizes up to the full 256^3 headline shape use size-independent properties only.
"""
import math

import numpy as np
import pytest

import oracle as O
import pica_amd as P
from helpers import courant_dt, plasma

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case] " + msg, flush=True)


def _run(ctx, prec, n, ppc, shape, dep, steps, seed=3):
    dt = courant_dt()
    n0, w0 = plasma(dt, ppc)
    ncell = (n, n, n)
    g = P.YeeGrid.periodic(ctx, ncell, prec=prec)
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), (4, 4, 4), prec=prec)
    ens.load_thermal(ppc, 0.01, 0.1 / math.sqrt(0.99), w0, seed)
    return dt, g, ens


def _continuity(ctx, prec, n, ppc, shape, dep):
    dt, g, ens = _run(ctx, prec, n, ppc, shape, dep, 0)
    P.run(ens, g, dt, 2, shape, dep)
    P.deposit_rho(ens, g, shape)
    rho0 = g.get(P.RHO).astype(np.float64)
    P.step(ens, g, dt, shape, dep)
    P.deposit_rho(ens, g, shape)
    rho1 = g.get(P.RHO).astype(np.float64)
    jx, jy, jz = (g.get(c).astype(np.float64) for c in (P.JX, P.JY, P.JZ))
    divj = O.divergence(jx, jy, jz, (1.0, 1.0, 1.0))
    drho = ((rho1 - rho0) / dt)[:-1, :-1, :-1]
    res = O.physical_nodes(drho + divj)
    scale = np.abs(O.physical_nodes(divj)).max()
    return float(np.abs(res).max() / scale)


@pytest.mark.parametrize("prec,tol,kernel", [(P.F64, 1e-11, P.KERNEL_AUTO), (P.F32, 2e-4, P.KERNEL_PER_PARTICLE),
                                             (P.F32, 2e-4, P.KERNEL_CELL_SORTED)])
@pytest.mark.parametrize("shape", [P.CIC, P.TSC])
def test_esirkepov_continuity(ctx, prec, tol, kernel, shape):
    _say("continuity prec=%d shape=%d kernel=%d" % (prec, shape, kernel))
    ctx.set_option(P.OPT_PARTICLE_KERNEL, kernel)
    try:
        r = _continuity(ctx, prec, 32, 8, shape, P.ESIRKEPOV)
    finally:
        ctx.set_option(P.OPT_PARTICLE_KERNEL, P.KERNEL_AUTO)
    print("continuity residual (prec=%d shape=%d kernel=%d): %.3g" % (prec, shape, kernel, r), flush=True)
    assert r < tol


def test_direct_deposition_is_not_charge_conserving(ctx):
    _say("continuity control (direct)")
    r = _continuity(ctx, P.F64, 32, 8, P.TSC, P.DIRECT)
    print("continuity residual of the direct depositor: %.3g" % r, flush=True)
    assert r > 1e-3


@pytest.mark.parametrize("prec,n,steps,tol", [(P.F64, 64, 20, 1e-11), (P.F32, 64, 20, 1e-4),
                                              (P.F32, 256, 5, 1e-4), (P.F64, 256, 5, 1e-11)])
def test_gauss_law_drift(ctx, prec, n, steps, tol):
    _say("gauss drift prec=%d n=%d" % (prec, n))
    dt, g, ens = _run(ctx, prec, n, 16 if n == 256 else 8, P.TSC, P.ESIRKEPOV, 0)
    P.diag_set_reference(ens, g, P.TSC)
    d0 = P.diag(ens, g, P.TSC)
    P.run(ens, g, dt, steps, P.TSC, P.ESIRKEPOV)
    d1 = P.diag(ens, g, P.TSC)
    drift = d1[P.DIAG_GAUSS_DRIFT] / d1[P.DIAG_RHO_MAX]
    q_rel = abs(d1[P.DIAG_CHARGE] - d0[P.DIAG_CHARGE]) / abs(d0[P.DIAG_CHARGE])
    print("gauss drift (prec=%d n=%d steps=%d): %.3g, charge change %.3g" % (prec, n, steps, drift, q_rel), flush=True)
    assert drift < tol
    assert q_rel < (1e-12 if prec == P.F64 else 1e-5)


@pytest.mark.parametrize("prec,tol", [(P.F64, 1e-11), (P.F32, 2e-4)])
@pytest.mark.parametrize("shape", [P.CIC, P.TSC])
def test_device_continuity_residual(ctx, prec, tol, shape):
    """pk_diag's PK_DIAG_CONTINUITY_MAX after pk_rho_save + one step equals the residual the host
    computes from the downloaded rho and J; Esirkepov keeps it at round-off, direct does not"""
    _say("device continuity prec=%d shape=%d" % (prec, shape))
    for dep in (P.ESIRKEPOV, P.DIRECT):
        dt, g, ens = _run(ctx, prec, 32, 8, shape, dep, 0)
        P.run(ens, g, dt, 2, shape, dep)
        P.deposit_rho(ens, g, shape)
        rho0 = g.get(P.RHO).astype(np.float64)
        P.rho_save(ens, g, shape)
        P.step(ens, g, dt, shape, dep)
        d = P.diag(ens, g, shape)
        P.deposit_rho(ens, g, shape)
        rho1 = g.get(P.RHO).astype(np.float64)
        jx, jy, jz = (g.get(c).astype(np.float64) for c in (P.JX, P.JY, P.JZ))
        divj = O.physical_nodes(O.divergence(jx, jy, jz, (1.0, 1.0, 1.0)))
        host = np.abs(O.physical_nodes(((rho1 - rho0) / dt)[:-1, :-1, :-1]) + divj).max()
        rel = d[P.DIAG_CONTINUITY_MAX] / d[P.DIAG_DIVJ_MAX]
        print("device continuity (prec=%d shape=%d dep=%d): %.3g (host %.3g), max|div J| %.3g vs host %.3g" %
              (prec, shape, dep, rel, host / np.abs(divj).max(), d[P.DIAG_DIVJ_MAX], np.abs(divj).max()), flush=True)
        assert d[P.DIAG_DIVJ_MAX] == pytest.approx(np.abs(divj).max(), rel=1e-6 if prec == P.F32 else 1e-12)
        if dep == P.ESIRKEPOV:
            assert rel < tol
        else:
            assert rel > 1e-3
            assert d[P.DIAG_CONTINUITY_MAX] == pytest.approx(host, rel=1e-3)
        ens.close()
        g.close()


@pytest.mark.parametrize("prec,tol", [(P.F64, 1e-11), (P.F32, 2e-4)])
def test_device_continuity_at_256cube(ctx, prec, tol):
    """the headline shape (256^3, 16 ppc, two-stream, TSC, Esirkepov): after three steps the device
    continuity residual of the next step, max |(rho1 - rho0)/dt + div J| over max |div J|, stays
    at round-off (a size-independent property; the oracle cannot run this size)"""
    _say("device continuity 256^3 prec=%d" % prec)
    dt, g, ens = _run(ctx, prec, 256, 16, P.TSC, P.ESIRKEPOV, 0)
    P.run(ens, g, dt, 3, P.TSC, P.ESIRKEPOV)
    P.rho_save(ens, g, P.TSC)
    P.step(ens, g, dt, P.TSC, P.ESIRKEPOV)
    d = P.diag(ens, g, P.TSC)
    r = d[P.DIAG_CONTINUITY_MAX] / d[P.DIAG_DIVJ_MAX]
    print("256^3 prec=%d: continuity residual %.3g of max |div J| %.3g" % (prec, r, d[P.DIAG_DIVJ_MAX]), flush=True)
    assert d[P.DIAG_DIVJ_MAX] > 0 and r < tol
    ens.close()
    g.close()
