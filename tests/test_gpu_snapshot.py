"""Snapshots (pk_grid_save / pk_grid_load / pk_ens_save / pk_ens_load) and the numpy reader
(pica_amd.snapshot): a load restores the saved state bit for bit and in the same particle order,
a run continued from it stays with the uninterrupted run (to the order of J sums), and the
reader sees exactly what the library holds."""
import math

import numpy as np
import pytest

import pica_amd as P
from helpers import courant_dt, plasma, rel_err
from pica_amd import snapshot as S

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case] " + msg, flush=True)


def _make(ctx, prec, ncell=(16, 8, 8), ppc=8):
    dt = courant_dt()
    _, w0 = plasma(dt, ppc)
    g = P.YeeGrid.periodic(ctx, ncell, prec=prec)
    e = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), (4, 4, 4), prec=prec)
    return dt, w0, g, e


@pytest.fixture
def period4(ctx):
    old = ctx.get_option(P.OPT_INTERLEAVE)
    ctx.set_option(P.OPT_INTERLEAVE, 4)
    yield
    ctx.set_option(P.OPT_INTERLEAVE, old)


@pytest.mark.parametrize("prec,tol", [(P.F64, 1e-12), (P.F32, 1e-5)])
def test_restart_from_snapshot(ctx, tmp_path, prec, tol, period4):
    _say("snapshot restart prec=%d" % prec)
    dt, w0, ga, ea = _make(ctx, prec)
    ea.load_thermal(8, 0.05, 0.1 / math.sqrt(0.99), w0, 77)
    # 3 + 5 steps with the interleave period at 4: the FP32 ensemble is reordered at re-sorts 4
    # and 8, which the continued run only matches if the load restores the re-sort count
    P.run(ea, ga, dt, 3, P.TSC, P.ESIRKEPOV)
    ea.save(tmp_path / "e.snap")
    ga.save(tmp_path / "g.snap")
    saved_p, saved_f = ea.particles(), ga.get_all()
    _, _, gb, eb = _make(ctx, prec)
    eb.load(tmp_path / "e.snap")
    gb.load(tmp_path / "g.snap")
    lp, lf = eb.particles(), gb.get_all()
    for a, b in zip(lp, saved_p):
        assert np.array_equal(a, b), "loaded particles differ from the saved ones"
    assert np.array_equal(lf, saved_f), "loaded fields differ from the saved ones"
    P.run(ea, ga, dt, 5, P.TSC, P.ESIRKEPOV)
    P.run(eb, gb, dt, 5, P.TSC, P.ESIRKEPOV)
    pa, pb = ea.particles(), eb.particles()
    assert np.array_equal(pa[3], pb[3]), "supercell membership differs after the continued run"
    errs = [rel_err(pb[0], pa[0]), rel_err(pb[1], pa[1])] + [rel_err(gb.get(c), ga.get(c)) for c in range(9)]
    print("continued run vs uninterrupted: max rel diff %.3g" % max(errs), flush=True)
    assert max(errs) < tol


@pytest.mark.parametrize("prec", [P.F64, P.F32])
def test_reader_sees_the_library_state(ctx, tmp_path, prec):
    _say("snapshot reader prec=%d" % prec)
    dt, w0, g, e = _make(ctx, prec, ncell=(8, 12, 4))
    e.load_thermal(4, 0.05, 0.0, w0, 5)
    P.run(e, g, dt, 2, P.CIC, P.ESIRKEPOV)
    e.save(tmp_path / "e.snap")
    g.save(tmp_path / "g.snap")
    x, u, w, off = e.particles()
    h, counts, fields = S.read_ensemble(tmp_path / "e.snap")
    assert h["prec"] == prec and int(h["count"]) == x.shape[0] and int(h["resorts"]) == 2
    assert np.array_equal(np.concatenate([[0], np.cumsum(counts)]), off)
    for d, name in enumerate("xyz"):
        assert np.array_equal(fields[name], x[:, d]), name
    for d, name in enumerate(("ux", "uy", "uz")):
        assert np.array_equal(fields[name].astype(np.float64), u[:, d]), name
    assert np.array_equal(fields["w"].astype(np.float64), w)
    hg, f = S.read_grid(tmp_path / "g.snap")
    for c in range(9):
        assert np.array_equal(f[c], g.get(c)), "component %d" % c


def test_load_rejects_other_geometry(ctx, tmp_path):
    _say("snapshot geometry check")
    dt, w0, g, e = _make(ctx, P.F64)
    e.load_thermal(2, 0.05, 0.0, w0, 3)
    e.save(tmp_path / "e.snap")
    g.save(tmp_path / "g.snap")
    other = P.EnsembleSupercells(ctx, (16, 8, 8), (1.0, 1.0, 1.0), (4, 4, 2))
    with pytest.raises(P.PicaError):
        other.load(tmp_path / "e.snap")
    g32 = P.YeeGrid.periodic(ctx, (16, 8, 8), prec=P.F32)
    with pytest.raises(P.PicaError):
        g32.load(tmp_path / "g.snap")
    with pytest.raises(P.PicaError):
        e.load(tmp_path / "g.snap")             # a grid file is not an ensemble
