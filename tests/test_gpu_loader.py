"""The device's seeded thermal loader (pk_ens_load_thermal) against its CPU twin in the oracle
(or_load_thermal): every particle bit-identical, in the same order, FP64 and FP32 split forms.
So CPU and GPU runs start from identical inputs by construction."""
import numpy as np
import pytest

import oracle as O
import pica_amd as P

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[case] " + msg, flush=True)


def _split_of(x, step, ncell):
    """the FP32 split form of absolute positions as the loader stores them (cell + float offset)"""
    c = np.empty(x.shape, np.int64)
    o = np.empty(x.shape, np.float32)
    for d in range(3):
        t = x[:, d] * (1.0 / step[d])
        cd = np.clip(np.floor(t).astype(np.int64), 0, ncell[d] - 1)
        f = (t - cd).astype(np.float32)
        f = np.where(f >= 0.0, f, np.float32(0.0))
        f = np.where(f >= 1.0, np.float32(np.nextafter(np.float32(1), np.float32(0))), f)
        c[:, d], o[:, d] = cd, f
    return c, o


@pytest.mark.parametrize("prec", [P.F64, P.F32])
@pytest.mark.parametrize("ppc,ncps", [(16, (4, 4, 4)), (5, (2, 4, 8))])
def test_device_loader_equals_cpu_twin(ctx, prec, ppc, ncps):
    _say("loader twin prec=%d ppc=%d ncps=%s" % (prec, ppc, ncps))
    ncell, step = (16, 16, 32), (1.0, 1.0, 1.0)
    uth, drift, w0, seed = 0.01, 0.1 / np.sqrt(0.99), 1.4e9, 20241203
    ens = P.EnsembleSupercells(ctx, ncell, step, ncps, prec=prec)
    ens.load_thermal(ppc, uth, drift, w0, seed)
    gx, gu, gw, goff = ens.particles()
    counts, gcell = ens.cells()
    x, u, w, cell = O.load_thermal(ncell, ncps, step, ppc, uth, drift, w0, seed, prec == P.F32)
    assert np.array_equal(gcell, cell), "cells differ"
    if prec == P.F64:
        assert np.array_equal(gx, x) and np.array_equal(gu, u) and np.array_equal(gw, w)
    else:
        c, o = _split_of(x, step, ncell)
        gc = np.stack([cell // (ncell[1] * ncell[2]), (cell // ncell[2]) % ncell[1], cell % ncell[2]], 1)
        o = np.where(c != gc, np.where(c < gc, np.float32(0), np.float32(np.nextafter(np.float32(1), np.float32(0)))), o)
        xa = (gc + o.astype(np.float64)) * np.array(step)
        assert np.array_equal(gx, xa), "positions differ"
        assert np.array_equal(gu, u.astype(np.float32).astype(np.float64)), "momenta differ"
        assert np.array_equal(gw, w.astype(np.float32).astype(np.float64))
    ens.close()
