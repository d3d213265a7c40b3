"""FP32 split-position form (int32 cell + float offset, FP32 fields/J) against the FP64 oracle.

The two runs start from identical seeded inputs.  Particles are matched through unique
weights (w_k = w0 (1 + k 2^-20), exact in FP32).  Tolerances are stated per quantity; they are
FP32 round-off bounds, not parity in the bit sense.
"""
import contextlib

import numpy as np
import pytest

import oracle as O
import pica_amd as P
from helpers import courant_dt, plasma, random_particles, rel_err

pytestmark = pytest.mark.gpu

# per-quantity bounds after 10 steps (measured errors are printed by the test)
TOL = {"x": 2e-6, "u": 2e-5, "E": 2e-4, "B": 2e-4, "J": 2e-4, "energy": 1e-5}


def _say(msg):
    print("[case] " + msg, flush=True)


def _inputs(seed, ncell, ppc, uth=0.05, drift=0.1):
    rng = np.random.default_rng(seed)
    dt = courant_dt()
    n0, w0 = plasma(dt, ppc)
    x, u = random_particles(rng, ncell, ppc, uth=uth, drift=drift)
    n = x.shape[0]
    w = (w0 * (1.0 + np.arange(n) * 2.0 ** -20)).astype(np.float32).astype(np.float64)
    # the FP32 form stores u and the in-cell offset in FP32: give the FP64 run the same values
    u = u.astype(np.float32).astype(np.float64)
    c = np.floor(x)
    x = c + (x - c).astype(np.float32).astype(np.float64)
    return dt, x, u, w


@contextlib.contextmanager
def _kernel(ctx, kernel, queue=0):
    """run with the given FP32 Esirkepov kernel and change-queue cap, restoring the defaults"""
    ctx.set_option(P.OPT_PARTICLE_KERNEL, kernel)
    ctx.set_option(P.OPT_CHANGE_QUEUE, queue)
    try:
        yield
    finally:
        ctx.set_option(P.OPT_PARTICLE_KERNEL, P.KERNEL_AUTO)
        ctx.set_option(P.OPT_CHANGE_QUEUE, 0)


def _compare(ctx, shape, dep, ncell, ncps, ppc, nsteps, seed, uth=0.05, drift=0.1, kernel=P.KERNEL_AUTO, queue=0):
    dt, x, u, w = _inputs(seed, ncell, ppc, uth, drift)
    og = O.Grid.periodic(ncell)
    oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    for _ in range(nsteps):
        O.step(og, oe, dt, shape, dep)
    e0 = sum(O.energy(og, oe))
    g = P.YeeGrid.periodic(ctx, ncell, prec=P.F32)
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps, prec=P.F32)
    ens.add(x, u, w)
    with _kernel(ctx, kernel, queue):
        P.run(ens, g, dt, nsteps, shape, dep)
    gx, gu, gw, goff = ens.particles()
    assert gx.shape[0] == oe.np
    d = P.diag(ens, g, shape)
    e1 = d[P.DIAG_FIELD_ENERGY] + d[P.DIAG_KINETIC_ENERGY]
    ia, ib = np.argsort(gw, kind="stable"), np.argsort(oe.w, kind="stable")
    assert np.array_equal(gw[ia], oe.w[ib]), "particle sets differ"
    L = np.array(ncell, np.float64)
    dx = gx[ia] - oe.positions()[ib]
    dx -= L * np.round(dx / L)                          # periodic images
    errs = {
        "x": float(np.abs(dx).max() / L.max()),
        "u": rel_err(gu[ia], oe.momenta()[ib]),
        "E": max(rel_err(g.get(c), og.f[c]) for c in (P.EX, P.EY, P.EZ)),
        "B": max(rel_err(g.get(c), og.f[c]) for c in (P.BX, P.BY, P.BZ)),
        "J": max(rel_err(g.get(c), og.f[c]) for c in (P.JX, P.JY, P.JZ)),
        "energy": abs(e1 - e0) / abs(e0),
    }
    print("split errors vs FP64 oracle (shape=%d dep=%d kernel=%d): %s" % (shape, dep, kernel, errs), flush=True)
    for k, v in errs.items():
        assert v < TOL[k], "%s error %.3g above %.3g" % (k, v, TOL[k])


@pytest.mark.parametrize("shape,dep,kernel", [(P.CIC, P.DIRECT, P.KERNEL_AUTO),
                                              (P.CIC, P.ESIRKEPOV, P.KERNEL_PER_PARTICLE),
                                              (P.CIC, P.ESIRKEPOV, P.KERNEL_CELL_SORTED),
                                              (P.TSC, P.DIRECT, P.KERNEL_AUTO),
                                              (P.TSC, P.ESIRKEPOV, P.KERNEL_PER_PARTICLE),
                                              (P.TSC, P.ESIRKEPOV, P.KERNEL_CELL_SORTED)])
def test_split_matches_fp64_oracle(ctx, shape, dep, kernel):
    _say("split vs fp64 shape=%d dep=%d kernel=%d" % (shape, dep, kernel))
    _compare(ctx, shape, dep, (16, 16, 16), (4, 4, 4), 8, 10, 41 + 2 * shape + dep, kernel=kernel)


@pytest.mark.parametrize("kernel", [P.KERNEL_PER_PARTICLE, P.KERNEL_CELL_SORTED])
def test_split_one_and_two_supercells_per_dimension(ctx, kernel):
    """8 x 4 x 12 cells in 4^3 supercells: 2, 1 and 3 supercells per dimension, so leavers wrap
    onto the neighbour from both sides (2) or back onto their own supercell (1)"""
    _say("split nsc=(2,1,3) kernel=%d" % kernel)
    _compare(ctx, P.TSC, P.ESIRKEPOV, (8, 4, 12), (4, 4, 4), 8, 10, 5, uth=0.1, kernel=kernel)


@pytest.mark.parametrize("ncps,ppc", [((2, 2, 2), 12), ((8, 8, 8), 6), ((4, 2, 8), 9)])
def test_split_supercell_shapes(ctx, ncps, ppc):
    """the two-particles-per-lane kernel with cubic supercells of 2 and 8 cells (their own
    compiled forms) and a non-cubic one (the general form), odd ppc included"""
    _say("split supercells %s ppc=%d" % (ncps, ppc))
    _compare(ctx, P.TSC, P.ESIRKEPOV, (16, 8, 16), ncps, ppc, 10, 31 + ppc, kernel=P.KERNEL_PER_PARTICLE)


@pytest.mark.parametrize("shape", [P.CIC, P.TSC])
def test_split_sorted_small_supercells(ctx, shape):
    """the cell-sorted kernel on 2^3 supercells"""
    _say("split sorted 2^3 shape=%d" % shape)
    _compare(ctx, shape, P.ESIRKEPOV, (8, 8, 6), (2, 2, 2), 12, 10, 9, kernel=P.KERNEL_CELL_SORTED)


def test_split_sorted_queue_overflow(ctx):
    """a hot plasma with the change queue capped at 8 entries: most cell changers deposit
    straight from the lane (largest window) instead of from the queue"""
    _say("split sorted queue cap 8")
    _compare(ctx, P.TSC, P.ESIRKEPOV, (12, 8, 8), (4, 4, 4), 10, 10, 13, uth=0.2, drift=0.0,
             kernel=P.KERNEL_CELL_SORTED, queue=8)


def test_split_cells_and_counts(ctx):
    """the cell of every particle and the per-cell counts match the FP64 oracle's positions"""
    _say("split cells")
    ncell, ncps, ppc = (8, 12, 8), (4, 4, 4), 6
    dt, x, u, w = _inputs(77, ncell, ppc)
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps, prec=P.F32)
    ens.add(x, u, w)
    counts, cells = ens.cells()
    gx, gu, gw, goff = ens.particles()
    oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    assert np.array_equal(goff, oe.off)
    ci = np.floor(oe.positions()).astype(np.int64)
    lin = (ci[:, 0] * ncell[1] + ci[:, 1]) * ncell[2] + ci[:, 2]
    assert np.array_equal(cells, lin)
    assert np.array_equal(counts.ravel(), np.bincount(lin, minlength=int(np.prod(ncell))))


def _interleave_order(cells, off, ncell, ncps):
    """the order pk_ens_interleave gives each supercell, with r the rank of a particle among those
    of its cell in the current order and h the cell's count: the pairs (ranks 2q, 2q + 1) by
    (q, local cell, r mod 2), then the last particle of each odd-count cell by local cell"""
    perm = np.arange(cells.shape[0])
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        c = cells[a:b]
        cz, cy, cx = c % ncell[2], (c // ncell[2]) % ncell[1], c // (ncell[1] * ncell[2])
        loc = ((cx % ncps[0]) * ncps[1] + cy % ncps[1]) * ncps[2] + cz % ncps[2]
        rank = np.zeros(b - a, np.int64)
        seen = {}
        for i, k in enumerate(loc):
            rank[i] = seen.get(k, 0)
            seen[k] = rank[i] + 1
        h = np.array([seen[k] for k in loc], np.int64)
        single = (h % 2 == 1) & (rank == h - 1)
        perm[a:b] = a + np.lexsort((rank % 2, loc, np.where(single, 0, rank // 2), single))
    return perm


@pytest.mark.parametrize("ppc", [6, 48])
def test_interleave_order_and_set(ctx, ppc):
    """pk_ens_interleave keeps every particle (bits included) and every supercell's membership,
    and orders each supercell as pairs of one cell by (pair rank, cell), odd leftovers last;
    48 ppc puts ~3072 particles in a supercell (the kernel form for up to 4096)"""
    _say("interleave ppc=%d" % ppc)
    ncell, ncps = (8, 12, 8), (4, 4, 4)
    dt, x, u, w = _inputs(88, ncell, ppc)
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps, prec=P.F32)
    g = P.YeeGrid.periodic(ctx, ncell, prec=P.F32)
    ens.add(x, u, w)
    period = ctx.get_option(P.OPT_INTERLEAVE)
    ctx.set_option(P.OPT_INTERLEAVE, 0)
    try:
        P.run(ens, g, dt, 6, P.TSC, P.ESIRKEPOV)          # let the order mix
    finally:
        ctx.set_option(P.OPT_INTERLEAVE, period)
    x0, u0, w0, off0 = ens.particles()
    _, cells0 = ens.cells()
    ens.interleave()
    x1, u1, w1, off1 = ens.particles()
    assert np.array_equal(off0, off1)
    perm = _interleave_order(cells0, off0, ncell, ncps)
    assert not np.array_equal(perm, np.arange(perm.shape[0]))
    assert np.array_equal(x1, x0[perm]) and np.array_equal(u1, u0[perm]) and np.array_equal(w1, w0[perm])
    _, cells1 = ens.cells()
    # the first pairs of each supercell now sit in distinct cells, two particles of one cell each
    for s in range(len(off1) - 1):
        head = cells1[off1[s]:off1[s + 1]][:32]
        assert np.array_equal(head[0::2], head[1::2])
        assert len(np.unique(head[0::2])) == len(head) // 2


@pytest.mark.parametrize("ppc", [4, 5])
def test_loader_pairs_hold_the_fp64_particles(ctx, ppc):
    """the FP32 loader places particle k of every cell in pairs (k = 2q, 2q + 1 adjacent, pairs
    by (q, cell), particle ppc - 1 of every cell last when ppc is odd); the FP64 loader places it
    at k * cells + cell; both hold the same particles (positions to FP32 rounding, momenta and
    weights exactly rounded)"""
    _say("loader pairs ppc=%d" % ppc)
    ncell, ncps = (8, 8, 12), (4, 4, 4)
    ens = {}
    for prec in (P.F64, P.F32):
        e = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps, prec=prec)
        e.load_thermal(ppc, 0.05, 0.1, 1.0, 9)
        ens[prec] = e
    x64, u64, w64, off = ens[P.F64].particles()
    x32, u32, w32, off32 = ens[P.F32].particles()
    assert np.array_equal(off, off32)
    ncsc = int(np.prod(ncps))
    npair = (ppc & ~1) * ncsc
    perm = []
    for s in range(len(off) - 1):
        for l in range(ppc * ncsc):
            if l < npair:
                lc, k = (l >> 1) % ncsc, 2 * ((l >> 1) // ncsc) + (l & 1)
            else:
                lc, k = l - npair, ppc - 1
            perm.append(off[s] + k * ncsc + lc)
    perm = np.array(perm)
    assert np.abs(x32 - x64[perm]).max() < 1e-5
    assert np.array_equal(u32, u64[perm].astype(np.float32).astype(np.float64))
    assert np.array_equal(w32, w64[perm].astype(np.float32).astype(np.float64))
    _, cells = ens[P.F32].cells()
    for s in range(len(off) - 1):
        c = cells[off[s]:off[s] + npair]
        assert np.array_equal(c[0::2], c[1::2])
    for e in ens.values():
        e.close()
