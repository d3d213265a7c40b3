"""The snapshot file format as documented (include/pica_amd.h, pica_amd/snapshot.py), written
here with numpy and read back by pica_amd.snapshot without the library or a GPU."""
import numpy as np
import pytest

from pica_amd import snapshot as S


def _header(kind, prec, esize, dims, nfield, count, nsct=0, ghost=0, ncps=(0, 0, 0), step=(1.0, 1.0, 1.0)):
    h = np.zeros(1, S.HEADER)
    h["magic"], h["kind"], h["version"], h["prec"] = b"PKSNAP1", kind, 1, prec
    h["dims"], h["ghost"], h["ncps"], h["nfield"] = dims, ghost, ncps, nfield
    h["step"], h["count"], h["nsct"], h["esize"] = step, count, nsct, esize
    return h.tobytes()


def test_grid_round_trip(tmp_path):
    dims = (5, 6, 7)
    rng = np.random.default_rng(1)
    f = rng.standard_normal((10,) + dims)
    p = tmp_path / "g.snap"
    p.write_bytes(_header(S.KIND_GRID, 0, 8, dims, 10, int(np.prod(dims)), ghost=2) + f.astype("<f8").tobytes())
    h, got = S.read_grid(p)
    assert h["ghost"] == 2 and np.array_equal(got, f)


def test_split_ensemble_positions(tmp_path):
    ncell, step = (4, 2, 3), (0.5, 1.0, 2.0)
    counts = np.array([3, 2], np.int32)
    cell = np.array([0, 5, 23, 7, 12], np.int32)
    off = np.array([0.25, 0.5, 0.75, 0.0, 0.999], np.float32)
    arrays = [cell] + [off] * 3 + [np.arange(5, dtype=np.float32)] * 3 + [np.ones(5, np.float32)]
    p = tmp_path / "e.snap"
    p.write_bytes(_header(S.KIND_ENSEMBLE, 1, 4, ncell, 8, 5, nsct=2, ncps=(2, 2, 3), step=step) +
                  counts.tobytes() + b"".join(a.tobytes() for a in arrays))
    h, c, fields = S.read_ensemble(p)
    assert np.array_equal(c, counts)
    cz = cell % 3
    cy = (cell // 3) % 2
    cx = cell // 6
    for name, cc, hd in (("x", cx, 0.5), ("y", cy, 1.0), ("z", cz, 2.0)):
        assert np.array_equal(fields[name], (cc + off.astype(np.float64)) * hd)


def test_rejects_other_files(tmp_path):
    p = tmp_path / "bad.snap"
    p.write_bytes(b"NOTASNAP" + bytes(152))
    with pytest.raises(ValueError):
        S.read_grid(p)
    q = tmp_path / "g.snap"
    q.write_bytes(_header(S.KIND_GRID, 0, 8, (5, 5, 5), 10, 125) + bytes(100))
    with pytest.raises(ValueError):
        S.read_ensemble(q)        # kind mismatch
    with pytest.raises(ValueError):
        S.read_grid(q)            # truncated
