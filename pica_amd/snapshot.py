"""Reader of the raw snapshot files written by pk_grid_save / pk_ens_save (numpy only, no GPU):
analysis, fixtures and training data away from the device.

File layout (little-endian): a 160-byte header, then raw arrays.
  header: magic "PKSNAP1\\0", kind (1 grid, 2 ensemble), version (1), prec (PK_F64 = 0,
          PK_F32 = 1), posmode, dims[3], ghost (grid) or species index (ensemble), ncps[3],
          nfield, origin[3], step[3], count, nsct, esize, a pad word, resorts (int64 at byte 128),
          6 reserved words;
  grid:     nfield components over the whole storage dims[0] x dims[1] x dims[2], z fastest;
  ensemble: nsct int32 supercell counts, then the fields in supercell order
            (F64: x y z ux uy uz w; F32 split: cell int32, ox oy oz ux uy uz w float32).
"""
import numpy as np

HEADER = np.dtype([
    ("magic", "S8"), ("kind", "<u4"), ("version", "<u4"), ("prec", "<u4"), ("posmode", "<u4"),
    ("dims", "<i4", (3,)), ("ghost", "<i4"), ("ncps", "<i4", (3,)), ("nfield", "<i4"),
    ("origin", "<f8", (3,)), ("step", "<f8", (3,)), ("count", "<i8"), ("nsct", "<i8"), ("esize", "<u4"),
    ("pad", "<u4"), ("resorts", "<i8"), ("reserved", "<u4", (6,)),
])
assert HEADER.itemsize == 160

KIND_GRID, KIND_ENSEMBLE = 1, 2
F64_FIELDS = ("x", "y", "z", "ux", "uy", "uz", "w")
SPLIT_FIELDS = ("cell", "ox", "oy", "oz", "ux", "uy", "uz", "w")


def _header(f, kind):
    h = np.frombuffer(f.read(HEADER.itemsize), HEADER, count=1)[0]
    if h["magic"] != b"PKSNAP1" or h["version"] != 1:
        raise ValueError("not a version-1 snapshot file")
    if h["kind"] != kind:
        raise ValueError("the snapshot holds kind %d, not %d" % (h["kind"], kind))
    return h


def _read(f, dtype, count):
    a = np.fromfile(f, dtype, count)
    if a.size != count:
        raise ValueError("truncated snapshot")
    return a


def read_grid(path):
    """(header, fields[nfield][dims])"""
    with open(path, "rb") as f:
        h = _header(f, KIND_GRID)
        dt = "<f8" if h["esize"] == 8 else "<f4"
        dims = tuple(int(v) for v in h["dims"])
        fields = np.stack([_read(f, dt, int(h["count"])).reshape(dims) for _ in range(int(h["nfield"]))])
    return h, fields


def read_ensemble(path):
    """(header, per-supercell counts, dict of field arrays in supercell order); split files also
    get absolute positions x, y, z = (cell coordinate + offset) * step"""
    with open(path, "rb") as f:
        h = _header(f, KIND_ENSEMBLE)
        nsct, n = int(h["nsct"]), int(h["count"])
        counts = _read(f, "<i4", nsct)
        out = {}
        if h["esize"] == 8:
            for name in F64_FIELDS:
                out[name] = _read(f, "<f8", n)
        else:
            for name in SPLIT_FIELDS:
                out[name] = _read(f, "<i4" if name == "cell" else "<f4", n)
            ncell = [int(v) for v in h["dims"]]
            c = out["cell"].astype(np.int64)
            coords = (c // (ncell[1] * ncell[2]), (c // ncell[2]) % ncell[1], c % ncell[2])
            for d, (name, o) in enumerate(zip("xyz", ("ox", "oy", "oz"))):
                out[name] = (coords[d] + out[o].astype(np.float64)) * h["step"][d]
    return h, counts, out
