"""The C++ surface (include/pica_amd/pica.hpp, pica's names over the C ABI) driven by a plain
g++ program, checked against the oracle: particles and fields of a short TSC/Esirkepov run,
restarted halfway from snapshot files and finished through the several-species step."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import pica_amd as P
from helpers import courant_dt, plasma, random_particles, rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cxx_surface_matches_oracle(tmp_path):
    print("[case] c++ surface", flush=True)
    exe = tmp_path / "surface_run"
    libdir = os.path.dirname(P.LIBPATH)
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cxx", "surface_run.cpp"), "-o", str(exe), "-L", libdir, "-lpica_amd",
                    "-Wl,-rpath," + libdir], check=True)
    rng = np.random.default_rng(17)
    ncell, ncps, ppc = (8, 8, 8), (4, 4, 4), 4
    dt = courant_dt()
    n0, w0 = plasma(dt, ppc)
    x, u = random_particles(rng, ncell, ppc, uth=0.05, drift=0.1)
    w = np.full(x.shape[0], w0)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array(ncell, np.int32).tofile(f)
        np.array(ncps, np.int32).tofile(f)
        np.array([x.shape[0]], np.int64).tofile(f)
        np.concatenate([x, u, w[:, None]], 1).astype(np.float64).tofile(f)
    out = tmp_path / "out.bin"
    r = subprocess.run([str(exe), str(inp), str(out), "5", "1", "1"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(out, np.uint8)
    n = int(raw[:8].view(np.int64)[0])
    p = raw[8:8 + n * 56].view(np.float64).reshape(n, 7)
    nsc = int(np.prod([a // b for a, b in zip(ncell, ncps)]))
    off = raw[8 + n * 56:8 + n * 56 + (nsc + 1) * 8].view(np.int64)
    fld = raw[8 + n * 56 + (nsc + 1) * 8:].view(np.float64).reshape((9,) + tuple(c + 4 for c in ncell))
    og = O.Grid.periodic(ncell)
    oe = O.Ensemble(ncell, ncps, (1.0, 1.0, 1.0), x, u, w)
    for _ in range(5):
        O.step(og, oe, dt, O.TSC, O.ESIRKEPOV)
    assert np.array_equal(off, oe.off)
    assert rel_err(p[:, :3], oe.positions()) < 1e-12
    assert rel_err(p[:, 3:6], oe.momenta()) < 1e-12
    for c in range(9):
        assert rel_err(fld[c], og.f[c]) < 1e-12


def test_pica_caller_switched_by_includes(tmp_path):
    """tests/cxx/pica_loop.cpp, a pica caller written against pica's API, built with
    -DUSE_PICA_AMD: only its includes and its particle loop differ (pica_compat.hpp supplies
    pica's YeeGrid<Three>, element accessors, Particle3d, ParticleArraySoA, EnsembleSupercells,
    YeeSolver, ParticleTypes).  Its result equals the oracle's CIC / direct run (particle order
    included: the library re-sorts as the oracle does) and, where oracle/_ref/pica_loop was built
    from pica's headers, the same program run on pica itself (as a set)."""
    print("[case] pica caller via pica_compat.hpp", flush=True)
    from pica_loop_io import case, oracle_run, read_output, write_input
    exe = tmp_path / "pica_loop"
    libdir = os.path.dirname(P.LIBPATH)
    subprocess.run(["g++", "-std=c++17", "-O2", "-DUSE_PICA_AMD", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cxx", "pica_loop.cpp"), "-o", str(exe), "-L", libdir, "-lpica_amd",
                    "-Wl,-rpath," + libdir], check=True)
    ncell, ncps, x, u, w, dt = case()
    write_input(tmp_path / "in.bin", ncell, ncps, x, u, w)
    r = subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "6"], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    p, f = read_output(tmp_path / "out.bin", ncell)
    og, oe = oracle_run(ncell, ncps, x, u, w, dt, 6)
    assert np.array_equal(p[:, 6], oe.w), "particle order differs from the oracle's"
    assert rel_err(p[:, :3], oe.positions()) < 1e-12 and rel_err(p[:, 3:6], oe.momenta()) < 1e-12
    for c in range(9):
        assert rel_err(f[c], og.f[c]) < 1e-12, "component %d" % c
    ref = os.path.join(ROOT, "oracle", "_ref", "pica_loop")
    if os.path.exists(ref):
        r = subprocess.run([ref, str(tmp_path / "in.bin"), str(tmp_path / "ref.bin"), "6"], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0
        q, g = read_output(tmp_path / "ref.bin", ncell)
        ia, ib = np.argsort(p[:, 6]), np.argsort(q[:, 6])
        assert np.array_equal(p[ia, 6], q[ib, 6])
        assert rel_err(p[ia, :6], q[ib, :6]) < 1e-12
        for c in range(9):
            assert rel_err(f[c], g[c]) < 1e-12, "component %d vs pica" % c
        print("pica_compat run equals pica's own run of the same program", flush=True)
