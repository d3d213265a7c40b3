"""The CPU oracle against golden fixtures written by pica's own code (tests/golden/make_golden.py).

These run anywhere (no pica sources, no GPU) and pin the oracle bit for bit: every fixture
output was produced by pica's headers compiled with the same -ffp-contract=off discipline.
"""
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "pica_units.npz"))


def test_yee_update_e_then_b_fp64():
    """YeeSolver::updateE then updateB on a 7x9x8 grid with unequal steps (YeeSolver.h:132-213)"""
    g = O.Grid(tuple(G["yee_s"]), tuple(G["yee_origin"]), tuple(G["yee_step"]), f=G["yee_in"].copy())
    O.yee_update_e(g, G["yee_dt"][0])
    assert np.array_equal(g.f, G["yee_after_e"])
    O.yee_update_b(g, G["yee_dt"][1])
    assert np.array_equal(g.f, G["yee_after_b"])


def test_yee_fp32():
    """YeeGrid<Three, float> (pica's float grid works, YeeSolverTests.cpp:30-42)"""
    g = O.Grid(tuple(G["yee_s"]), tuple(G["yee_origin"]), tuple(G["yee_step"]), np.float32, G["yee32_in"].copy())
    O.yee_update_e(g, np.float32(1.1e-11))
    O.yee_update_b(g, np.float32(0.7e-11))
    assert np.array_equal(g.f, G["yee32_out"])


def test_interp_cic_plain():
    """FieldInterpolatorCIC<YeeGrid<Three>>::get (FieldInterpolatorYeeGrid.h:150-182)"""
    g = O.Grid((12, 12, 12), (-2.0, -2.0, -2.0), (1.0, 1.0, 1.0), f=G["interp_f"])
    e, b = O.interp_cic(g, G["interp_pos"])
    assert np.array_equal(e, G["interp_e"]) and np.array_equal(b, G["interp_b"])


def test_interp_cic_supercell():
    """FieldInterpolatorCICSupercell (FieldInterpolatorYeeGrid.h:197-283)"""
    g = O.Grid((12, 12, 12), (-2.0, -2.0, -2.0), (1.0, 1.0, 1.0), f=G["interp_f"])
    e, b = O.interp_sc(g, tuple(G["isc_min"]), (4, 4, 4), O.CIC, G["isc_pos"])
    assert np.array_equal(e, G["isc_e"]) and np.array_equal(b, G["isc_b"])


def test_deposit_cic_plain():
    """CurrentDepositorCIC::deposit, no 1/V (CurrentDepositorYeeGrid.h:138-163)"""
    g = O.Grid((12, 12, 12), (-2.0, -2.0, -2.0), (1.0, 1.0, 1.0))
    O.deposit_cic(g, G["interp_pos"], G["dep_cur"])
    assert np.array_equal(g.f[6:9], G["dep_j"])


def test_deposit_cic_supercell():
    """CurrentDepositorCICSupercell: tile, flush x 1/V on destruction (CurrentDepositorYeeGrid.h:178-262)"""
    g = O.Grid((12, 12, 12), (-2.0, -2.0, -2.0), (1.0, 1.0, 1.0))
    O.deposit_direct_sc(g, tuple(G["isc_min"]), (4, 4, 4), O.CIC, G["dsc_pos"], G["dsc_cur"])
    assert np.array_equal(g.f[6:9], G["dsc_j"])


def test_boris_push():
    """BorisPusher::push on ParticleArraySoA (BorisPusher.h:32-47, ParticleArray.h:209,257-263)"""
    x, u = O.boris_push(G["boris_x"], G["boris_u"], G["boris_e"], G["boris_b"], O.E_CHARGE, O.E_MASS,
                        G["boris_dt"][0])
    assert np.array_equal(x, G["boris_x_out"]) and np.array_equal(u, G["boris_u_out"])


def test_cell_and_supercell_indices():
    """Grid::getCellIndex truncates, EnsembleSupercells::getSupercellIndex floors (Grid.h:36,
    EnsembleSupercells.h:53-57); includes pica's (5.2, 7.5, -1.4) vector (VectorsTests.cpp:426-436)"""
    g = O.Grid((20, 20, 20), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    assert np.array_equal(O.cell_index_grid(g, G["idx_pos"]), G["idx_cell"])
    si = O.supercell_index((0.0, 0.0, 0.0), (16.0, 16.0, 16.0), (16, 16, 16), (4, 2, 1), G["idx_pos"])
    assert np.array_equal(si, G["idx_sc"])
    k = np.where(np.all(G["idx_pos"] == [5.2, 7.5, -1.4], axis=1))[0][0]
    assert tuple(G["idx_cell"][k]) == (5, 7, -1)
    sk = O.supercell_index((0.0, 0.0, 0.0), (16.0, 16.0, 16.0), (16, 16, 16), (1, 1, 1), G["idx_pos"][k:k + 1])
    assert tuple(sk[0]) == (5, 7, -2)


def test_small_periodic_run():
    """6 periodic steps of pica's supercell interpolator + Boris + supercell CIC depositor + Yee"""
    ens = O.Ensemble((8, 8, 12), (4, 4, 4), (1.0, 1.0, 1.0), G["run_x"], G["run_u"], G["run_w"])
    g = O.Grid.periodic((8, 8, 12))
    g.f[...] = G["run_f"]
    dt = 0.5 / O.C
    for _ in range(6):
        O.step(g, ens, dt, O.CIC, O.DIRECT)
    assert np.array_equal(ens.off, G["run_off"])
    assert np.array_equal(ens.positions(), G["run_x_out"])
    assert np.array_equal(ens.momenta(), G["run_u_out"])
    assert np.array_equal(ens.w, G["run_w_out"])
    assert np.array_equal(g.f, G["run_f_out"])


@pytest.mark.parametrize("nsteps", [50, 100])
def test_thermal32_summary(nsteps):
    """the thermal 32^3 plasma (8 ppc, CIC, direct, FP64): membership, per-cell counts, sampled particles
    and field planes after 50 and 100 steps equal pica's"""
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_golden import thermal32_inputs
    S = np.load(os.path.join(HERE, "golden", "thermal32_cic_direct_summary.npz"))
    ncell, dt, x, u, w = thermal32_inputs(int(S["seed"][0]))
    ens = O.Ensemble(ncell, (4, 4, 4), (1.0, 1.0, 1.0), x, u, w)
    g = O.Grid.periodic(ncell)
    for _ in range(nsteps):
        O.step(g, ens, dt, O.CIC, O.DIRECT)
    assert np.array_equal(ens.off, S["off_%d" % nsteps])
    ci = np.floor(ens.positions()).astype(np.int64)
    counts = np.bincount((ci[:, 0] * 32 + ci[:, 1]) * 32 + ci[:, 2], minlength=32 ** 3)
    assert np.array_equal(counts, S["counts_%d" % nsteps])
    sample = np.arange(0, ens.np, 263)
    assert np.array_equal(ens.positions()[sample], S["sample_x_%d" % nsteps])
    assert np.array_equal(ens.momenta()[sample], S["sample_u_%d" % nsteps])
    phys = g.f[:, 2:-2, 2:-2, 2:-2]
    assert np.array_equal(phys[:, 5], S["plane_%d" % nsteps])
    assert np.array_equal((phys ** 2).sum(axis=(1, 2, 3)), S["sumsq_%d" % nsteps])
