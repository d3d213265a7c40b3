"""CPU oracle for the pica PIC step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package, and only as the checker.  The product (``pica_amd``) never imports it.

``liboracle.so`` (built from ``pica_oracle.c``) restates pica's arithmetic operation by
operation (see the header of pica_oracle.c for the file:line map).  ``_ref/libpica_ref.so`` is
pica's own header-only code compiled through ``ref_harness.cpp``; it exists only where the
pica sources were present at build time and is used to pin the restatement.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

EX, EY, EZ, BX, BY, BZ, JX, JY, JZ = range(9)
CIC, TSC = 0, 1
DIRECT, ESIRKEPOV = 0, 1
PI = 3.14159265358          # Constants.h:10
C = 29979245800.0           # Constants.h:11
E_CHARGE = -4.80320427e-10  # Constants.h:12
E_MASS = 9.10938215e-28     # Constants.h:13
GHOST = 2

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


class _GridD(ctypes.Structure):
    _fields_ = [("s", ctypes.c_int * 3), ("origin", ctypes.c_double * 3), ("step", ctypes.c_double * 3),
                ("f", _dp * 9)]


class _GridF(ctypes.Structure):
    _fields_ = [("s", ctypes.c_int * 3), ("origin", ctypes.c_float * 3), ("step", ctypes.c_float * 3),
                ("f", _fp * 9)]


class _Ens(ctypes.Structure):
    _fields_ = [("nsc", ctypes.c_int * 3), ("ncps", ctypes.c_int * 3), ("ncell", ctypes.c_int * 3),
                ("ghost", ctypes.c_int), ("step", ctypes.c_double * 3), ("np", ctypes.c_long),
                ("x", _dp * 3), ("u", _dp * 3), ("w", _dp), ("off", ctypes.POINTER(ctypes.c_long)),
                ("cls", ctypes.POINTER(ctypes.c_byte)), ("q", ctypes.c_double), ("m", ctypes.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _LIB = ctypes.CDLL(path)
        _LIB.or_boris_coeff.restype = ctypes.c_double
    return _LIB


def ref():
    """pica's own code (oracle/_ref/libpica_ref.so) or None where it was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libpica_ref.so")
        if not os.path.exists(path):
            return None
        _REF = ctypes.CDLL(path)
    return _REF


def _ptr(a, ctype=ctypes.c_double):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class Grid:
    """A YeeGrid<Three, Real>: nine z-fastest arrays of shape s (incl. ghosts)."""

    def __init__(self, s, origin, step, dtype=np.float64, f=None):
        self.s = tuple(int(v) for v in s)
        self.origin = tuple(float(v) for v in origin)
        self.step = tuple(float(v) for v in step)
        self.dtype = np.dtype(dtype)
        self.f = np.zeros((9,) + self.s, dtype=self.dtype) if f is None else np.ascontiguousarray(f, self.dtype)

    @classmethod
    def periodic(cls, ncell, step=(1.0, 1.0, 1.0), dtype=np.float64):
        s = tuple(n + 2 * GHOST for n in ncell)
        origin = tuple(-GHOST * h for h in step)
        return cls(s, origin, step, dtype)

    def copy(self):
        return Grid(self.s, self.origin, self.step, self.dtype, self.f.copy())

    def c(self):
        if self.dtype == np.float64:
            g = _GridD()
            for d in range(3):
                g.s[d], g.origin[d], g.step[d] = self.s[d], self.origin[d], self.step[d]
            for c in range(9):
                g.f[c] = _ptr(self.f[c])
        else:
            g = _GridF()
            for d in range(3):
                g.s[d], g.origin[d], g.step[d] = self.s[d], self.origin[d], self.step[d]
            for c in range(9):
                g.f[c] = _ptr(self.f[c], ctypes.c_float)
        self._keep = g
        return ctypes.byref(g)

    def comps_ptrs(self):
        ct = ctypes.c_double if self.dtype == np.float64 else ctypes.c_float
        arr = (ctypes.POINTER(ct) * 9)(*[_ptr(self.f[c], ct) for c in range(9)])
        self._keep_p = arr
        return arr

    def physical(self, comp):
        g = GHOST
        return self.f[comp][g:-g, g:-g, g:-g]


def yee_update_e(grid, dt):
    fn = lib().or_yee_update_e if grid.dtype == np.float64 else lib().or_yee_update_e_f32
    ct = ctypes.c_double if grid.dtype == np.float64 else ctypes.c_float
    fn(grid.c(), ct(dt))


def yee_update_b(grid, dt):
    fn = lib().or_yee_update_b if grid.dtype == np.float64 else lib().or_yee_update_b_f32
    ct = ctypes.c_double if grid.dtype == np.float64 else ctypes.c_float
    fn(grid.c(), ct(dt))


def ghost_fill(arr, G=GHOST):
    s = (ctypes.c_int * 3)(*arr.shape)
    if arr.dtype == np.float64:
        lib().or_ghost_fill(s, G, _ptr(arr))
    else:
        lib().or_ghost_fill_f32(s, G, _ptr(arr, ctypes.c_float))


def ghost_fold(arr, G=GHOST):
    s = (ctypes.c_int * 3)(*arr.shape)
    if arr.dtype == np.float64:
        lib().or_ghost_fold(s, G, _ptr(arr))
    else:
        lib().or_ghost_fold_f32(s, G, _ptr(arr, ctypes.c_float))


def _xyz(x):
    x = [np.ascontiguousarray(x[:, d], np.float64) for d in range(3)]
    return x, (_dp * 3)(*[_ptr(v) for v in x])


def interp_cic(grid, pos):
    """pica FieldInterpolatorCIC::get for positions pos (n x 3); returns e, b (n x 3)."""
    n = pos.shape[0]
    xs, xp = _xyz(pos)
    e = [np.zeros(n) for _ in range(3)]
    b = [np.zeros(n) for _ in range(3)]
    lib().or_interp_cic(grid.c(), ctypes.c_long(n), xp, (_dp * 3)(*[_ptr(v) for v in e]),
                        (_dp * 3)(*[_ptr(v) for v in b]))
    return np.stack(e, 1), np.stack(b, 1)


def deposit_cic(grid, pos, cur):
    n = pos.shape[0]
    xs, xp = _xyz(pos)
    cs, cp = _xyz(cur)
    lib().or_deposit_cic(grid.c(), ctypes.c_long(n), xp, cp)


def boris_coeff(q, m, dt):
    return lib().or_boris_coeff(ctypes.c_double(q), ctypes.c_double(m), ctypes.c_double(dt))


def boris_push(x, u, e, b, q, m, dt):
    """pica BorisPusher::push on n particles; returns new (x, u)."""
    n = x.shape[0]
    xs, xp = _xyz(x)
    us, up = _xyz(u)
    es, ep = _xyz(e)
    bs, bp = _xyz(b)
    lib().or_boris_push(ctypes.c_long(n), ctypes.c_double(boris_coeff(q, m, dt)), ctypes.c_double(dt), xp, up,
                        ep, bp)
    return np.stack(xs, 1), np.stack(us, 1)


def cell_index_grid(grid, pos):
    n = pos.shape[0]
    xs, xp = _xyz(pos)
    out = np.zeros((n, 3), np.int32)
    lib().or_cell_index_grid(grid.c(), ctypes.c_long(n), xp, _ptr(out, ctypes.c_int))
    return out


def supercell_index(mn, mx, ncell, ncps, pos):
    n = pos.shape[0]
    xs, xp = _xyz(pos)
    out = np.zeros((n, 3), np.int32)
    lib().or_supercell_index((ctypes.c_double * 3)(*mn), (ctypes.c_double * 3)(*mx), (ctypes.c_int * 3)(*ncell),
                             (ctypes.c_int * 3)(*ncps), ctypes.c_long(n), xp, _ptr(out, ctypes.c_int))
    return out


class Ensemble:
    """Supercell-ordered ensemble of one species (positions absolute, domain [0, ncell*step))."""

    def __init__(self, ncell, ncps, step, x, u, w, q=E_CHARGE, m=E_MASS, bucket=True):
        self.ncell = tuple(int(v) for v in ncell)
        self.ncps = tuple(int(v) for v in ncps)
        self.step = tuple(float(v) for v in step)
        self.nsc = tuple((a + b - 1) // b for a, b in zip(self.ncell, self.ncps))
        self.q, self.m = float(q), float(m)
        self.x = np.ascontiguousarray(np.asarray(x, np.float64).T).copy()   # 3 x n
        self.u = np.ascontiguousarray(np.asarray(u, np.float64).T).copy()
        self.w = np.ascontiguousarray(np.asarray(w, np.float64)).copy()
        self.np = self.w.shape[0]
        self.off = np.zeros(int(np.prod(self.nsc)) + 1, np.int64)
        self.cls = np.full(max(self.np, 1), 13, np.int8)
        if bucket:
            r = lib().or_bucket(self.c())
            if r:
                raise ValueError("particle outside the domain")

    def copy(self):
        e = Ensemble.__new__(Ensemble)
        e.__dict__.update({k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()
                           if not k.startswith("_")})
        return e

    def c(self):
        e = _Ens()
        for d in range(3):
            e.nsc[d], e.ncps[d], e.ncell[d], e.step[d] = self.nsc[d], self.ncps[d], self.ncell[d], self.step[d]
        e.ghost = GHOST
        e.np = self.np
        for d in range(3):
            e.x[d] = _ptr(self.x[d])
            e.u[d] = _ptr(self.u[d])
        e.w = _ptr(self.w)
        e.off = _ptr(self.off, ctypes.c_long)
        e.cls = _ptr(self.cls, ctypes.c_byte)
        e.q, e.m = self.q, self.m
        self._keep = e
        return ctypes.byref(e)

    def positions(self):
        return self.x.T.copy()

    def momenta(self):
        return self.u.T.copy()

    def counts(self):
        return np.diff(self.off)


def particle_stage(grid, ens, dt, shape=CIC, dep=DIRECT, nthreads=0):
    r = lib().or_particle_stage(grid.c(), ens.c(), ctypes.c_double(dt), shape, dep, nthreads)
    if r:
        raise RuntimeError("or_particle_stage failed: %d" % r)


def field_stage(grid, dt):
    lib().or_field_stage(grid.c(), GHOST, ctypes.c_double(dt))


def resort(ens):
    if lib().or_resort(ens.c()):
        raise RuntimeError("or_resort failed")


def step(grid, ens, dt, shape=CIC, dep=DIRECT, nthreads=0):
    r = lib().or_step(grid.c(), ens.c(), ctypes.c_double(dt), shape, dep, nthreads)
    if r:
        raise RuntimeError("or_step failed: %d" % r)


def deposit_rho(grid, ens, shape=CIC):
    rho = np.zeros(grid.s, np.float64)
    if lib().or_deposit_rho(grid.c(), ens.c(), shape, _ptr(rho)):
        raise RuntimeError("or_deposit_rho: particle outside the grid")
    return rho


def energy(grid, ens):
    out = (ctypes.c_double * 2)()
    lib().or_energy(grid.c(), ens.c(), out)
    return out[0], out[1]


# ---------------------------------------------------------------------------------------------
# Diagnostics on the Yee grid (numpy).  rho nodes are cell centres; E/J_x sits on x faces.
# ---------------------------------------------------------------------------------------------
def divergence(fx, fy, fz, step):
    """div at cell centres of face-centred vector components (E or J): sum_d (F_d(a+e_d) - F_d(a)) / h_d.
    Returns an array of the interior shape s - 1 (index a of the result = node a)."""
    d = (fx[1:, :-1, :-1] - fx[:-1, :-1, :-1]) / step[0]
    d = d + (fy[:-1, 1:, :-1] - fy[:-1, :-1, :-1]) / step[1]
    d = d + (fz[:-1, :-1, 1:] - fz[:-1, :-1, :-1]) / step[2]
    return d


def physical_nodes(arr):
    """physical cell-centre nodes of a divergence array (indices GHOST..GHOST+N-1)."""
    g = GHOST
    return arr[g:g + (arr.shape[0] + 1 - 2 * g), g:g + (arr.shape[1] + 1 - 2 * g), g:g + (arr.shape[2] + 1 - 2 * g)]


def ref_run_cic_direct(ncell, step, ncps, q, m, dt, nsteps, x, u, w, f=None, nthreads=0):
    """pica's own supercell path, nsteps periodic steps (see ref_harness.cpp)."""
    r = ref()
    if r is None:
        raise RuntimeError("pica reference harness not built")
    n = x.shape[0]
    xs = [np.ascontiguousarray(x[:, d], np.float64).copy() for d in range(3)]
    us = [np.ascontiguousarray(u[:, d], np.float64).copy() for d in range(3)]
    ws = np.ascontiguousarray(w, np.float64).copy()
    nsc = [(a + b - 1) // b for a, b in zip(ncell, ncps)]
    off = np.zeros(int(np.prod(nsc)) + 1, np.int64)
    grid = Grid.periodic(ncell, step)
    if f is not None:
        grid.f[...] = f
    rc = r.ref_run_cic_direct((ctypes.c_int * 3)(*ncell), (ctypes.c_double * 3)(*step), (ctypes.c_int * 3)(*ncps),
                              ctypes.c_double(q), ctypes.c_double(m), ctypes.c_double(dt), nsteps, ctypes.c_long(n),
                              _ptr(xs[0]), _ptr(xs[1]), _ptr(xs[2]), _ptr(us[0]), _ptr(us[1]), _ptr(us[2]),
                              _ptr(ws), _ptr(off, ctypes.c_long), grid.comps_ptrs(), nthreads)
    if rc:
        raise RuntimeError("ref_run_cic_direct failed: %d" % rc)
    return np.stack(xs, 1), np.stack(us, 1), ws, off, grid


def interp_sc(grid, scmin, ncps, shape, pos):
    """supercell interpolation (pica's FieldInterpolatorCICSupercell for CIC) of positions pos."""
    n = pos.shape[0]
    xs, xp = _xyz(pos)
    e = [np.zeros(n) for _ in range(3)]
    b = [np.zeros(n) for _ in range(3)]
    r = lib().or_interp_sc(grid.c(), (ctypes.c_double * 3)(*scmin), (ctypes.c_int * 3)(*ncps), shape,
                           ctypes.c_long(n), xp, (_dp * 3)(*[_ptr(v) for v in e]), (_dp * 3)(*[_ptr(v) for v in b]))
    if r:
        raise ValueError("or_interp_sc: position outside the supercell tile (%d)" % r)
    return np.stack(e, 1), np.stack(b, 1)


def deposit_direct_sc(grid, scmin, ncps, shape, pos, cur):
    """supercell direct deposition (pica's CurrentDepositorCICSupercell for CIC), flushed x 1/V."""
    n = pos.shape[0]
    xs, xp = _xyz(pos)
    cs, cp = _xyz(cur)
    r = lib().or_deposit_direct_sc(grid.c(), (ctypes.c_double * 3)(*scmin), (ctypes.c_int * 3)(*ncps), shape,
                                   ctypes.c_long(n), xp, cp)
    if r:
        raise ValueError("or_deposit_direct_sc: position outside the supercell tile (%d)" % r)


def deposit_esirkepov_sc(grid, scmin, ncps, shape, x0, x1, qw, dt):
    """Esirkepov deposition of particles moving x0 -> x1 with charges qw, flushed x 1/V."""
    n = x0.shape[0]
    a, ap = _xyz(x0)
    b, bp = _xyz(x1)
    q = np.ascontiguousarray(qw, np.float64)
    r = lib().or_deposit_esirkepov_sc(grid.c(), (ctypes.c_double * 3)(*scmin), (ctypes.c_int * 3)(*ncps), shape,
                                      ctypes.c_long(n), ap, bp, _ptr(q), ctypes.c_double(dt))
    if r:
        raise ValueError("or_deposit_esirkepov_sc: move outside the supercell tile (%d)" % r)


class _Ref:
    """thin access to pica's own entry points in oracle/_ref (see ref_harness.cpp)"""

    @staticmethod
    def yee(kind, grid, dt):
        r = ref()
        s = (ctypes.c_int * 3)(*grid.s)
        if grid.dtype == np.float64:
            fn = r.ref_yee_update_e if kind == "e" else r.ref_yee_update_b
            fn(s, (ctypes.c_double * 3)(*grid.origin), (ctypes.c_double * 3)(*grid.step), grid.comps_ptrs(),
               ctypes.c_double(dt))
        else:
            fn = r.ref_yee_update_e_f32 if kind == "e" else r.ref_yee_update_b_f32
            fn(s, (ctypes.c_float * 3)(*grid.origin), (ctypes.c_float * 3)(*grid.step), grid.comps_ptrs(),
               ctypes.c_float(dt))

    @staticmethod
    def interp_cic(grid, pos, scmin=None, ncps=None):
        r = ref()
        n = pos.shape[0]
        xs = [np.ascontiguousarray(pos[:, d], np.float64) for d in range(3)]
        e = np.zeros((n, 3))
        b = np.zeros((n, 3))
        args = [(ctypes.c_int * 3)(*grid.s), (ctypes.c_double * 3)(*grid.origin), (ctypes.c_double * 3)(*grid.step),
                grid.comps_ptrs()]
        if scmin is None:
            r.ref_interp_cic(*args, ctypes.c_long(n), _ptr(xs[0]), _ptr(xs[1]), _ptr(xs[2]), _ptr(e), _ptr(b))
        else:
            r.ref_interp_cic_sc(*args, (ctypes.c_double * 3)(*scmin), (ctypes.c_int * 3)(*ncps), ctypes.c_long(n),
                                _ptr(xs[0]), _ptr(xs[1]), _ptr(xs[2]), _ptr(e), _ptr(b))
        return e, b

    @staticmethod
    def deposit_cic(grid, pos, cur, scmin=None, ncps=None):
        r = ref()
        n = pos.shape[0]
        xs = [np.ascontiguousarray(pos[:, d], np.float64) for d in range(3)]
        c = np.ascontiguousarray(cur, np.float64)
        args = [(ctypes.c_int * 3)(*grid.s), (ctypes.c_double * 3)(*grid.origin), (ctypes.c_double * 3)(*grid.step),
                grid.comps_ptrs()]
        if scmin is None:
            r.ref_deposit_cic(*args, ctypes.c_long(n), _ptr(xs[0]), _ptr(xs[1]), _ptr(xs[2]), _ptr(c))
        else:
            r.ref_deposit_cic_sc(*args, (ctypes.c_double * 3)(*scmin), (ctypes.c_int * 3)(*ncps), ctypes.c_long(n),
                                 _ptr(xs[0]), _ptr(xs[1]), _ptr(xs[2]), _ptr(c))

    @staticmethod
    def boris(x, u, e, b, w, q, m, dt):
        r = ref()
        n = x.shape[0]
        xx = [np.ascontiguousarray(x[:, d], np.float64).copy() for d in range(3)]
        uu = [np.ascontiguousarray(u[:, d], np.float64).copy() for d in range(3)]
        r.ref_boris_push(ctypes.c_long(n), ctypes.c_double(q), ctypes.c_double(m), ctypes.c_double(dt), _ptr(xx[0]),
                         _ptr(xx[1]), _ptr(xx[2]), _ptr(uu[0]), _ptr(uu[1]), _ptr(uu[2]),
                         _ptr(np.ascontiguousarray(w, np.float64)), _ptr(np.ascontiguousarray(e, np.float64)),
                         _ptr(np.ascontiguousarray(b, np.float64)))
        return np.stack(xx, 1), np.stack(uu, 1)

    @staticmethod
    def cell_index(grid, pos):
        r = ref()
        n = pos.shape[0]
        xs = [np.ascontiguousarray(pos[:, d], np.float64) for d in range(3)]
        out = np.zeros((n, 3), np.int32)
        r.ref_cell_index((ctypes.c_int * 3)(*grid.s), (ctypes.c_double * 3)(*grid.origin),
                         (ctypes.c_double * 3)(*grid.step), ctypes.c_long(n), _ptr(xs[0]), _ptr(xs[1]), _ptr(xs[2]),
                         _ptr(out, ctypes.c_int))
        return out

    @staticmethod
    def supercell_index(mn, mx, ncell, ncps, pos):
        r = ref()
        n = pos.shape[0]
        xs = [np.ascontiguousarray(pos[:, d], np.float64) for d in range(3)]
        out = np.zeros((n, 3), np.int32)
        r.ref_supercell_index((ctypes.c_double * 3)(*mn), (ctypes.c_double * 3)(*mx), (ctypes.c_int * 3)(*ncell),
                              (ctypes.c_int * 3)(*ncps), ctypes.c_long(n), _ptr(xs[0]), _ptr(xs[1]), _ptr(xs[2]),
                              _ptr(out, ctypes.c_int))
        return out


REF = _Ref()


def load_thermal(ncell, ncps, step, ppc, uth, drift, weight, seed, paired):
    """the library's seeded thermal loader restated on the CPU: particles in the order the device
    stores them (supercell by supercell, slot order; paired = the FP32 split form's order).
    Returns x (n, 3), u (n, 3), w (n,), cell (n,) -- cell is the linear physical cell, z fastest."""
    nsc = [a // b for a, b in zip(ncell, ncps)]
    n = int(np.prod(nsc)) * int(np.prod(ncps)) * int(ppc)
    x = [np.zeros(n) for _ in range(3)]
    u = [np.zeros(n) for _ in range(3)]
    w = np.zeros(n)
    cell = np.zeros(n, np.int32)
    lib().or_load_thermal((ctypes.c_int * 3)(*ncell), (ctypes.c_int * 3)(*ncps), (ctypes.c_double * 3)(*step),
                          int(ppc), ctypes.c_double(uth), ctypes.c_double(drift), ctypes.c_double(weight),
                          ctypes.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), int(bool(paired)),
                          (_dp * 3)(*[_ptr(v) for v in x]), (_dp * 3)(*[_ptr(v) for v in u]), _ptr(w),
                          _ptr(cell, ctypes.c_int))
    return np.stack(x, 1), np.stack(u, 1), w, cell


def det_log(x):
    """the loader's ln(x), 0 < x <= 1"""
    f = lib().or_det_log_x
    f.restype = ctypes.c_double
    return f(ctypes.c_double(x))


def det_sincos_2pi(U):
    """the loader's (sin, cos)(2 pi U), 0 <= U < 1"""
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().or_det_sincos_2pi_x(ctypes.c_double(U), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value
