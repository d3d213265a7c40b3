"""tests/cxx/pica_loop.cpp built against pica's own headers (oracle/_ref/pica_loop, where the pica
sources are present) is a genuine pica caller: its loop over supercells of
FieldInterpolatorCICSupercell, BorisPusher and CurrentDepositorCICSupercell, with its own
migration and periodic ghosts, matches the oracle's CIC / direct step (particles matched as a set
through unique weights: the caller's migration orders them its own way)."""
import os
import subprocess

import numpy as np
import pytest

from helpers import rel_err
from pica_loop_io import case, oracle_run, read_output, write_input

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "pica_loop")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="pica sources not present (oracle/_ref not built)")


def test_pica_caller_loop_matches_oracle(tmp_path):
    ncell, ncps, x, u, w, dt = case()
    write_input(tmp_path / "in.bin", ncell, ncps, x, u, w)
    r = subprocess.run([EXE, str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "6"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    p, f = read_output(tmp_path / "out.bin", ncell)
    og, oe = oracle_run(ncell, ncps, x, u, w, dt, 6)
    ia, ib = np.argsort(p[:, 6]), np.argsort(oe.w)
    assert np.array_equal(p[ia, 6], oe.w[ib])
    assert rel_err(p[ia, :3], oe.positions()[ib]) < 1e-12
    assert rel_err(p[ia, 3:6], oe.momenta()[ib]) < 1e-12
    for c in range(9):
        assert rel_err(f[c], og.f[c]) < 1e-12, "component %d" % c
