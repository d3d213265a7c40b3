#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from pica's own code.

pica is header-only C++; oracle/ref_harness.cpp compiles its headers (from the reference
checkout, never copied here) into oracle/_ref/libpica_ref.so, and this script calls it on
seeded inputs.  The fixtures hold inputs (or the seed that regenerates them) and pica's
outputs; tests/test_golden.py checks the CPU oracle against them where the pica sources are
absent (e.g. on a machine that only has this repository).

    make -C oracle && python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

THERMAL32_SEED = 20240601
SMALL_SEED = 1234


def _cd3(v):
    return (ctypes.c_double * 3)(*v)


def _ci3(v):
    return (ctypes.c_int * 3)(*v)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def thermal32_inputs(seed=THERMAL32_SEED):
    """32^3 thermal plasma, 8 ppc, sigma_u = 0.01, omega_p dt = 0.1, Courant 0.5 (CGS, h = 1 cm)."""
    rng = np.random.default_rng(seed)
    ncell, ppc = (32, 32, 32), 8
    dt = 0.5 / O.C
    wp = 0.1 / dt
    n0 = wp * wp * O.E_MASS / (4.0 * O.PI * O.E_CHARGE * O.E_CHARGE)
    n = 32 ** 3 * ppc
    cells = np.repeat(np.arange(32 ** 3), ppc)
    x = np.stack([cells // 1024 + rng.random(n), (cells // 32) % 32 + rng.random(n), cells % 32 + rng.random(n)], 1)
    u = rng.standard_normal((n, 3)) * 0.01
    w = np.full(n, n0 / ppc)
    return ncell, dt, x.astype(np.float64), u, w


def small_run_inputs(seed=SMALL_SEED):
    rng = np.random.default_rng(seed)
    ncell, ppc = (8, 8, 12), 4
    dt = 0.5 / O.C
    n = int(np.prod(ncell)) * ppc
    cells = np.repeat(np.arange(int(np.prod(ncell))), ppc)
    x = np.stack([cells // 96 + rng.random(n), (cells // 12) % 8 + rng.random(n), cells % 12 + rng.random(n)], 1)
    u = rng.standard_normal((n, 3)) * 0.08
    w = rng.random(n) * 1e9 + 1e9
    f = np.zeros((9, 12, 12, 16))
    f[:6] = rng.standard_normal((6, 12, 12, 16)) * 3e-3
    for c in range(6):
        O.ghost_fill(f[c])
    return ncell, dt, x, u, w, f


def main():
    R = O.ref()
    if R is None:
        sys.exit("pica reference harness not built (make -C oracle with the pica sources present)")
    rng = np.random.default_rng(99)
    out = {}

    # --- Yee solver, FP64 and FP32, on a non-cubic grid with unequal steps
    s, origin, step = (7, 9, 8), (-2.0, -1.0, -3.0), (1.0, 0.5, 2.0)
    f0 = rng.standard_normal((9,) + s)
    g = O.Grid(s, origin, step, f=f0.copy())
    R.ref_yee_update_e(_ci3(s), _cd3(origin), _cd3(step), g.comps_ptrs(), ctypes.c_double(1.1e-11))
    fe = g.f.copy()
    R.ref_yee_update_b(_ci3(s), _cd3(origin), _cd3(step), g.comps_ptrs(), ctypes.c_double(0.7e-11))
    out.update(yee_in=f0, yee_after_e=fe, yee_after_b=g.f.copy(), yee_s=np.array(s), yee_origin=np.array(origin),
               yee_step=np.array(step), yee_dt=np.array([1.1e-11, 0.7e-11]))
    f0f = rng.standard_normal((9,) + s).astype(np.float32)
    gf = O.Grid(s, origin, step, np.float32, f0f.copy())
    R.ref_yee_update_e_f32(_ci3(s), (ctypes.c_float * 3)(*origin), (ctypes.c_float * 3)(*step), gf.comps_ptrs(),
                           ctypes.c_float(1.1e-11))
    R.ref_yee_update_b_f32(_ci3(s), (ctypes.c_float * 3)(*origin), (ctypes.c_float * 3)(*step), gf.comps_ptrs(),
                           ctypes.c_float(0.7e-11))
    out.update(yee32_in=f0f, yee32_out=gf.f.copy())

    # --- CIC interpolation and deposition, plain (whole grid) and supercell variants
    s, origin, step = (12, 12, 12), (-2.0, -2.0, -2.0), (1.0, 1.0, 1.0)
    fi = rng.standard_normal((9,) + s)
    pos = rng.random((300, 3)) * 7.0 + 0.5
    xs = [np.ascontiguousarray(pos[:, d]) for d in range(3)]
    e = np.zeros((300, 3))
    b = np.zeros((300, 3))
    gi = O.Grid(s, origin, step, f=fi.copy())
    R.ref_interp_cic(_ci3(s), _cd3(origin), _cd3(step), gi.comps_ptrs(), ctypes.c_long(300), _dp(xs[0]), _dp(xs[1]),
                     _dp(xs[2]), _dp(e), _dp(b))
    out.update(interp_f=fi, interp_pos=pos, interp_e=e.copy(), interp_b=b.copy())
    scmin, ncps = (4.0, 4.0, 0.0), (4, 4, 4)
    psc = np.array(scmin) + rng.random((200, 3)) * 4.0
    xs = [np.ascontiguousarray(psc[:, d]) for d in range(3)]
    e2 = np.zeros((200, 3))
    b2 = np.zeros((200, 3))
    R.ref_interp_cic_sc(_ci3(s), _cd3(origin), _cd3(step), gi.comps_ptrs(), _cd3(scmin), _ci3(ncps), ctypes.c_long(200),
                        _dp(xs[0]), _dp(xs[1]), _dp(xs[2]), _dp(e2), _dp(b2))
    out.update(isc_min=np.array(scmin), isc_pos=psc, isc_e=e2, isc_b=b2)
    cur = rng.standard_normal((300, 3)) * 1e8
    gd = O.Grid(s, origin, step)
    xs = [np.ascontiguousarray(pos[:, d]) for d in range(3)]
    R.ref_deposit_cic(_ci3(s), _cd3(origin), _cd3(step), gd.comps_ptrs(), ctypes.c_long(300), _dp(xs[0]), _dp(xs[1]),
                      _dp(xs[2]), _dp(np.ascontiguousarray(cur)))
    out.update(dep_cur=cur, dep_j=gd.f[6:9].copy())
    # supercell depositor: positions in [P0 - h/2, P0 + (n+1) h), the range its tile covers
    pds = np.array(scmin) - 0.5 + rng.random((200, 3)) * 5.0
    curs = rng.standard_normal((200, 3)) * 1e8
    gds = O.Grid(s, origin, step)
    xs = [np.ascontiguousarray(pds[:, d]) for d in range(3)]
    R.ref_deposit_cic_sc(_ci3(s), _cd3(origin), _cd3(step), gds.comps_ptrs(), _cd3(scmin), _ci3(ncps),
                         ctypes.c_long(200), _dp(xs[0]), _dp(xs[1]), _dp(xs[2]), _dp(np.ascontiguousarray(curs)))
    out.update(dsc_pos=pds, dsc_cur=curs, dsc_j=gds.f[6:9].copy())

    # --- Boris pusher
    n = 400
    bx = rng.random((n, 3)) * 10.0
    bu = rng.standard_normal((n, 3)) * np.array([0.01, 1.0, 30.0])[rng.integers(0, 3, n)][:, None]
    be = rng.standard_normal((n, 3)) * 5.0
    bb = rng.standard_normal((n, 3)) * 50.0
    bw = rng.random(n) + 0.5
    xx = [np.ascontiguousarray(bx[:, d]).copy() for d in range(3)]
    uu = [np.ascontiguousarray(bu[:, d]).copy() for d in range(3)]
    dtb = 1.3e-12
    R.ref_boris_push(ctypes.c_long(n), ctypes.c_double(O.E_CHARGE), ctypes.c_double(O.E_MASS), ctypes.c_double(dtb),
                     _dp(xx[0]), _dp(xx[1]), _dp(xx[2]), _dp(uu[0]), _dp(uu[1]), _dp(uu[2]), _dp(bw),
                     _dp(np.ascontiguousarray(be)), _dp(np.ascontiguousarray(bb)))
    out.update(boris_x=bx, boris_u=bu, boris_e=be, boris_b=bb, boris_w=bw, boris_dt=np.array([dtb]),
               boris_x_out=np.stack(xx, 1), boris_u_out=np.stack(uu, 1))

    # --- indices: Grid::getCellIndex (truncate) and EnsembleSupercells::getSupercellIndex (floor)
    ip = np.concatenate([rng.random((300, 3)) * 20.0 - 2.0,
                         np.array([[5.2, 7.5, -1.4], [0.0, 4.0, 8.0], [3.999999999, 4.000000001, -0.0]])])
    xs = [np.ascontiguousarray(ip[:, d]) for d in range(3)]
    ci = np.zeros((ip.shape[0], 3), np.int32)
    R.ref_cell_index(_ci3((20, 20, 20)), _cd3((0.0, 0.0, 0.0)), _cd3((1.0, 1.0, 1.0)), ctypes.c_long(ip.shape[0]),
                     _dp(xs[0]), _dp(xs[1]), _dp(xs[2]), ci.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    si = np.zeros((ip.shape[0], 3), np.int32)
    R.ref_supercell_index(_cd3((0.0, 0.0, 0.0)), _cd3((16.0, 16.0, 16.0)), _ci3((16, 16, 16)), _ci3((4, 2, 1)),
                          ctypes.c_long(ip.shape[0]), _dp(xs[0]), _dp(xs[1]), _dp(xs[2]),
                          si.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    out.update(idx_pos=ip, idx_cell=ci, idx_sc=si)

    # --- a small periodic run, 6 steps, random initial fields (all inputs stored)
    ncell, dt, x, u, w, f = small_run_inputs()
    xr, ur, wr, offr, gr = O.ref_run_cic_direct(ncell, (1.0, 1.0, 1.0), (4, 4, 4), O.E_CHARGE, O.E_MASS, dt, 6, x, u,
                                                w, f)
    out.update(run_x=x, run_u=u, run_w=w, run_f=f, run_x_out=xr, run_u_out=ur, run_w_out=wr, run_off=offr,
               run_f_out=gr.f)
    np.savez_compressed(os.path.join(HERE, "pica_units.npz"), **out)

    # --- the thermal 32^3 plasma, 8 ppc, CIC, direct, 100 steps: a summary (inputs regenerate from the seed)
    ncell, dt, x, u, w = thermal32_inputs()
    summ = {"seed": np.array([THERMAL32_SEED])}
    for nsteps in (50, 100):
        xr, ur, wr, offr, gr = O.ref_run_cic_direct(ncell, (1.0, 1.0, 1.0), (4, 4, 4), O.E_CHARGE, O.E_MASS, dt,
                                                    nsteps, x, u, w)
        sample = np.arange(0, xr.shape[0], 263)
        ci = np.floor(xr).astype(np.int64)
        counts = np.bincount((ci[:, 0] * 32 + ci[:, 1]) * 32 + ci[:, 2], minlength=32 ** 3).astype(np.int16)
        g = O.GHOST
        phys = gr.f[:, g:-g, g:-g, g:-g]
        summ.update({"off_%d" % nsteps: offr, "counts_%d" % nsteps: counts, "sample_x_%d" % nsteps: xr[sample],
                     "sample_u_%d" % nsteps: ur[sample], "sumsq_%d" % nsteps: (phys ** 2).sum(axis=(1, 2, 3)),
                     "plane_%d" % nsteps: phys[:, 5].copy()})
    np.savez_compressed(os.path.join(HERE, "thermal32_cic_direct_summary.npz"), **summ)
    print("wrote", os.path.join(HERE, "pica_units.npz"), "and thermal32_cic_direct_summary.npz")


if __name__ == "__main__":
    main()
