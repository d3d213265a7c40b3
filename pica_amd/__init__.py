"""pica_amd -- the pica particle-in-cell step on MI355X (gfx950), behind pica's own names.

The compute lives in ``libpica_amd.so`` (hand-written HIP, built in-tree by
``pica_amd/build.py``); this module is a thin ctypes layer over its C ABI
(``include/pica_amd.h``) with pica's type and method names:

    YeeGrid(ctx, size, origin, step)          pica::YeeGrid<Three, Real>   (YeeGrid.h:191-256)
    YeeSolver().updateE(grid, dt)/updateB     pica::YeeSolver              (YeeSolver.h:11-38)
    EnsembleSupercells(ctx, ncell, step, ncps) pica::EnsembleSupercells    (EnsembleSupercells.h:15-69)
    particle_step(ens, grid, dt, shape, dep)  FieldInterpolatorCICSupercell::get + BorisPusher::push +
                                              CurrentDepositorCICSupercell::deposit for every particle

There is no CPU fallback: without the native library every entry point raises.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libpica_amd.so")

PK_OK, PK_E_ARG, PK_E_HIP, PK_E_ALLOC, PK_E_STATE, PK_E_UNSUPPORTED, PK_E_CAPACITY = 0, -1, -2, -3, -4, -5, -6
F64, F32 = 0, 1
POS_ABS, POS_SPLIT = 0, 1
CIC, TSC = 0, 1
DIRECT, ESIRKEPOV = 0, 1
EX, EY, EZ, BX, BY, BZ, JX, JY, JZ, RHO = range(10)
H2D, D2H = 0, 1
FILL_E, FILL_B, FILL_J, FOLD_J, FOLD_RHO = 1, 2, 4, 8, 16
(DIAG_FIELD_ENERGY, DIAG_KINETIC_ENERGY, DIAG_CHARGE, DIAG_GAUSS_MAX, DIAG_GAUSS_DRIFT, DIAG_RHO_MAX,
 DIAG_CONTINUITY_MAX, DIAG_DIVJ_MAX) = range(8)
DIAG_N = 8
OPT_PARTICLE_KERNEL, OPT_CHANGE_QUEUE, OPT_INTERLEAVE, OPT_INTERLEAVE_F64 = 0, 1, 2, 3
BATCH_WG_PER_CU, BATCH_ORDER, BATCH_GRAPH, BATCH_SUPERCELL, BATCH_INTERLEAVE, BATCH_BLOCK = 0, 1, 2, 3, 4, 5
SHAPE_PER_INSTANCE = 2
KERNEL_AUTO, KERNEL_PER_PARTICLE, KERNEL_CELL_SORTED = 0, 1, 2

# pica's CGS constants (Constants.h:10-13)
PI = 3.14159265358
C = 29979245800.0
ELECTRON_CHARGE = -4.80320427e-10
ELECTRON_MASS = 9.10938215e-28

# every symbol include/pica_amd.h declares
ABI_SYMBOLS = (
    "pk_ctx_create", "pk_ctx_destroy", "pk_last_error", "pk_sync", "pk_ctx_stream", "pk_timer_enable",
    "pk_timer_get", "pk_timer_reset", "pk_ctx_set_option", "pk_ctx_get_option", "pk_species_set",
    "pk_grid_create", "pk_grid_destroy", "pk_grid_copy",
    "pk_grid_zero", "pk_yee_update_e", "pk_yee_update_b", "pk_ghost", "pk_ens_create", "pk_ens_destroy",
    "pk_ens_set_slack", "pk_ens_upload", "pk_ens_download", "pk_ens_size", "pk_ens_num_supercells",
    "pk_ens_cells", "pk_ens_load_thermal", "pk_particle_step", "pk_ens_resort", "pk_step", "pk_run",
    "pk_deposit_rho", "pk_diag", "pk_diag_set_reference", "pk_batch_create", "pk_batch_destroy", "pk_batch_set",
    "pk_batch_get", "pk_batch_run", "pk_batch_autotune", "pk_ens_load_modes", "pk_dn_create", "pk_dn_destroy",
    "pk_dn_num_params", "pk_dn_init", "pk_dn_set_params", "pk_dn_get_params", "pk_dn_set_scale", "pk_dn_forward_host",
    "pk_dn_layer_host", "pk_dn_grad_host", "pk_dn_grad_batch_host", "pk_dn_add_pair", "pk_dn_num_pairs",
    "pk_dn_clear_pairs", "pk_dn_train", "pk_dn_apply", "pk_rho_save", "pk_project_continuity", "pk_step_denoised", "pk_grid_save", "pk_grid_load",
    "pk_ens_save", "pk_ens_load", "pk_step_species", "pk_ens_interleave", "pk_diag_energies_static",
    "pk_ens_set_supercells", "pk_batch_set_shape", "pk_dn_set_charge_gain", "pk_dn_add_pair_crops",
    "pk_run_captured",
)

_LIB = None
_CONTEXTS = weakref.WeakSet()


@atexit.register
def _close_all():
    # release device objects before the interpreter (and then the HIP runtime) shut down
    for c in list(_CONTEXTS):
        try:
            c.close()
        except Exception:
            pass


class PicaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pica_amd error %d: %s" % (code, msg))
        self.code = code


def lib():
    """The native library; raises if it has not been built (there is no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("pica_amd: native library %s is missing; run `python -m pica_amd.build`" % LIBPATH)
        L = ctypes.CDLL(LIBPATH)
        vp = ctypes.c_void_p
        L.pk_last_error.restype = ctypes.c_char_p
        L.pk_last_error.argtypes = [vp]
        L.pk_ctx_stream.restype = vp
        L.pk_ctx_stream.argtypes = [vp]
        L.pk_ctx_destroy.argtypes = [vp]
        L.pk_grid_destroy.argtypes = [vp]
        L.pk_ens_destroy.argtypes = [vp]
        L.pk_ens_size.restype = ctypes.c_int64
        L.pk_ens_size.argtypes = [vp]
        L.pk_ens_num_supercells.restype = ctypes.c_int64
        L.pk_ens_num_supercells.argtypes = [vp]
        for name in ("pk_sync", "pk_timer_reset"):
            getattr(L, name).argtypes = [vp]
        L.pk_timer_enable.argtypes = [vp, ctypes.c_int]
        L.pk_ctx_set_option.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.pk_ctx_get_option.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.pk_timer_get.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]
        L.pk_yee_update_e.argtypes = [vp, ctypes.c_double]
        L.pk_yee_update_b.argtypes = [vp, ctypes.c_double]
        L.pk_ghost.argtypes = [vp, ctypes.c_int]
        L.pk_grid_copy.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
        L.pk_grid_zero.argtypes = [vp, ctypes.c_int]
        L.pk_ens_set_slack.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.pk_ens_upload.argtypes = [vp, ctypes.c_size_t] + [vp] * 7
        L.pk_ens_download.argtypes = [vp] * 9
        L.pk_ens_cells.argtypes = [vp, vp, vp]
        L.pk_ens_load_thermal.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_uint64]
        L.pk_particle_step.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.pk_ens_resort.argtypes = [vp]
        L.pk_step.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.pk_run.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.pk_run_captured.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.pk_deposit_rho.argtypes = [vp, vp, ctypes.c_int]
        L.pk_diag.argtypes = [vp, vp, ctypes.c_int, vp]
        L.pk_diag_set_reference.argtypes = [vp, vp, ctypes.c_int]
        L.pk_diag_energies_static.argtypes = [vp, vp, vp]
        L.pk_ens_load_modes.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_uint64, ctypes.c_int, vp]
        L.pk_dn_create.argtypes = [vp, ctypes.POINTER(vp)]
        L.pk_dn_destroy.argtypes = [vp]
        L.pk_dn_num_params.argtypes = []
        for name in ("pk_dn_set_params", "pk_dn_get_params"):
            getattr(L, name).argtypes = [vp, vp]
        L.pk_dn_set_scale.argtypes = [vp, ctypes.c_double]
        L.pk_dn_set_charge_gain.argtypes = [vp, ctypes.c_int, ctypes.c_double, vp]
        L.pk_dn_init.argtypes = [vp, ctypes.c_uint64]
        L.pk_dn_forward_host.argtypes = [vp, vp, vp, vp]
        L.pk_dn_layer_host.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.pk_dn_grad_host.argtypes = [vp, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_double)]
        L.pk_dn_grad_batch_host.argtypes = [vp, vp, ctypes.c_int, vp, vp, vp, ctypes.POINTER(ctypes.c_double)]
        L.pk_dn_add_pair.argtypes = [vp, vp, vp]
        L.pk_dn_add_pair_crops.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, vp]
        L.pk_dn_num_pairs.argtypes = [vp]
        L.pk_dn_clear_pairs.argtypes = [vp]
        L.pk_dn_train.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, vp]
        L.pk_dn_apply.argtypes = [vp, vp]
        L.pk_rho_save.argtypes = [vp, vp, ctypes.c_int]
        L.pk_step_species.argtypes = [vp, ctypes.c_int, vp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.pk_ens_interleave.argtypes = [vp]
        L.pk_ens_set_supercells.argtypes = [vp, vp]
        L.pk_batch_set_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        for name in ("pk_grid_save", "pk_grid_load", "pk_ens_save", "pk_ens_load"):
            getattr(L, name).argtypes = [vp, ctypes.c_char_p]
        L.pk_project_continuity.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double]
        L.pk_step_denoised.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.pk_batch_create.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.POINTER(vp)]
        L.pk_batch_destroy.argtypes = [vp]
        L.pk_batch_set.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.pk_batch_get.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.pk_batch_run.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.pk_batch_autotune.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double)]
        _LIB = L
    return _LIB


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _d3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


class Context:
    """One device and one HIP stream (pk_ctx).  Grids and ensembles created on a context are
    released before it (close() and interpreter exit)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        r = lib().pk_ctx_create(int(device), ctypes.byref(h))
        if r:
            raise PicaError(r, "pk_ctx_create(device=%d) failed" % device)
        self.h = h
        self._children = weakref.WeakSet()
        _CONTEXTS.add(self)

    def _adopt(self, obj):
        self._children.add(obj)

    def check(self, r):
        if r:
            raise PicaError(r, lib().pk_last_error(self.h).decode())

    def sync(self):
        self.check(lib().pk_sync(self.h))

    def stream(self):
        return lib().pk_ctx_stream(self.h)

    def set_species(self, mass, charge):
        m = np.ascontiguousarray(mass, np.float64)
        q = np.ascontiguousarray(charge, np.float64)
        self.check(lib().pk_species_set(self.h, len(m), _p(m), _p(q)))

    def timer_enable(self, on=True):
        self.check(lib().pk_timer_enable(self.h, 1 if on else 0))

    def timer_reset(self):
        self.check(lib().pk_timer_reset(self.h))

    def set_option(self, option, value):
        """pk_ctx_set_option, e.g. (OPT_PARTICLE_KERNEL, KERNEL_CELL_SORTED)"""
        self.check(lib().pk_ctx_set_option(self.h, int(option), int(value)))

    def get_option(self, option):
        v = ctypes.c_int()
        self.check(lib().pk_ctx_get_option(self.h, int(option), ctypes.byref(v)))
        return v.value

    def timer(self, kernel):
        ms, n = ctypes.c_double(), ctypes.c_long()
        self.check(lib().pk_timer_get(self.h, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "h", None):
            for child in list(self._children):
                child.close()
            lib().pk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class YeeGrid:
    """pica::YeeGrid<Three, Real>(origin, step, size) on the device (+ rho), z-fastest arrays."""

    def __init__(self, ctx, size, origin, step, prec=F64, ghost=0):
        self.ctx = ctx
        self.size = tuple(int(v) for v in size)
        self.origin = tuple(float(v) for v in origin)
        self.step = tuple(float(v) for v in step)
        self.prec = prec
        self.ghost = ghost
        self.dtype = np.float64 if prec == F64 else np.float32
        h = ctypes.c_void_p()
        ctx.check(lib().pk_grid_create(ctx.h, prec, _i3(self.size), _d3(self.origin), _d3(self.step), int(ghost),
                                       ctypes.byref(h)))
        self.h = h
        ctx._adopt(self)

    @classmethod
    def periodic(cls, ctx, ncell, step=(1.0, 1.0, 1.0), prec=F64):
        """ncell physical cells with 2 periodic ghost layers per side, origin -2 step (the
        convention pica's supercell interpolator and depositor assume)."""
        return cls(ctx, [n + 4 for n in ncell], [-2.0 * h for h in step], step, prec, ghost=2)

    def get(self, comp):
        a = np.empty(self.size, self.dtype)
        self.ctx.check(lib().pk_grid_copy(self.h, comp, _p(a), D2H))
        return a

    def set(self, comp, arr):
        a = np.ascontiguousarray(arr, self.dtype)
        assert a.shape == self.size
        self.ctx.check(lib().pk_grid_copy(self.h, comp, _p(a), H2D))

    def get_all(self):
        return np.stack([self.get(c) for c in range(9)])

    def set_all(self, f):
        for c in range(9):
            self.set(c, f[c])

    def zero(self, comp):
        self.ctx.check(lib().pk_grid_zero(self.h, comp))

    def ghost_ops(self, ops):
        self.ctx.check(lib().pk_ghost(self.h, ops))

    def save(self, path):
        """raw snapshot of every component (pica_amd.snapshot reads it without a GPU)"""
        self.ctx.check(lib().pk_grid_save(self.h, os.fsencode(path)))

    def load(self, path):
        """restore a snapshot of a grid of the same geometry and precision, bit for bit"""
        self.ctx.check(lib().pk_grid_load(self.h, os.fsencode(path)))

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                lib().pk_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class YeeSolver:
    """pica::YeeSolver (YeeSolver.h:11-38): stateless, pica's method names."""

    def updateE(self, grid, dt):
        grid.ctx.check(lib().pk_yee_update_e(grid.h, float(dt)))

    def updateB(self, grid, dt):
        grid.ctx.check(lib().pk_yee_update_b(grid.h, float(dt)))


class EnsembleSupercells:
    """pica::EnsembleSupercells over ParticleArraySoA on the device, one species per ensemble.
    The domain is [0, ncell*step); ncell must be a multiple of ncps."""

    def __init__(self, ctx, ncell, step, ncps, prec=F64, posmode=None, type_index=0):
        self.ctx = ctx
        self.ncell = tuple(int(v) for v in ncell)
        self.step = tuple(float(v) for v in step)
        self.ncps = tuple(int(v) for v in ncps)
        self.prec = prec
        self.posmode = (POS_ABS if prec == F64 else POS_SPLIT) if posmode is None else posmode
        h = ctypes.c_void_p()
        ctx.check(lib().pk_ens_create(ctx.h, prec, self.posmode, _i3(self.ncell), _d3(self.step), _i3(self.ncps),
                                      int(type_index), ctypes.byref(h)))
        self.h = h
        ctx._adopt(self)

    def set_slack(self, frac, add):
        self.ctx.check(lib().pk_ens_set_slack(self.h, float(frac), int(add)))

    def add(self, x, u, w):
        """EnsembleSupercells::add for every row of x (n x 3 positions), u (n x 3, p/(mc)), w."""
        x = np.asarray(x, np.float64)
        u = np.asarray(u, np.float64)
        if x.ndim != 2 or x.shape[1] != 3:
            raise ValueError("add: x must be an (n, 3) array of positions")
        n = x.shape[0]
        if u.shape != (n, 3):
            raise ValueError("add: u must have the shape of x, (%d, 3); got %s" % (n, u.shape))
        w = np.asarray(w, np.float64)
        if w.ndim == 0:
            w = np.full(n, float(w))
        if w.shape != (n,):
            raise ValueError("add: w must be a scalar or hold one weight per particle (%d); got %s" % (n, w.shape))
        cols = [np.ascontiguousarray(x[:, d]) for d in range(3)] + [np.ascontiguousarray(u[:, d]) for d in range(3)]
        cols.append(np.ascontiguousarray(w))
        self.ctx.check(lib().pk_ens_upload(self.h, x.shape[0], *[_p(c) for c in cols]))

    def load_thermal(self, ppc, uth, drift_x, weight, seed):
        self.ctx.check(lib().pk_ens_load_thermal(self.h, int(ppc), float(uth), float(drift_x), float(weight),
                                                 int(seed) & 0xFFFFFFFFFFFFFFFF))

    def load_modes(self, ppc, uth, drift_x, weight, seed, modes):
        """load_thermal plus velocity modes: rows (kx, ky, kz, ax, ay, az, phase),
        u += a sin(2 pi k . x / L + phase)"""
        m = np.ascontiguousarray(modes, np.float64).reshape(-1, 7)
        self.ctx.check(lib().pk_ens_load_modes(self.h, int(ppc), float(uth), float(drift_x), float(weight),
                                               int(seed) & 0xFFFFFFFFFFFFFFFF, m.shape[0], _p(m)))

    def size(self):
        n = lib().pk_ens_size(self.h)
        if n < 0:
            self.ctx.check(PK_E_STATE)
        return int(n)

    def num_supercells(self):
        return int(lib().pk_ens_num_supercells(self.h))

    def particles(self):
        """(x n x 3, u n x 3, w n, off nsc+1) in supercell order."""
        n = self.size()
        cols = [np.empty(n) for _ in range(7)]
        off = np.empty(self.num_supercells() + 1, np.int64)
        self.ctx.check(lib().pk_ens_download(self.h, *[_p(c) for c in cols], _p(off)))
        return np.stack(cols[:3], 1), np.stack(cols[3:6], 1), cols[6], off

    def cells(self):
        """(per-cell counts over the physical cells, linear cell index of every particle)."""
        n = self.size()
        counts = np.empty(int(np.prod(self.ncell)), np.int32)
        cells = np.empty(max(n, 1), np.int32)
        self.ctx.check(lib().pk_ens_cells(self.h, _p(counts), _p(cells)))
        return counts.reshape(self.ncell), cells[:n]

    def resort(self):
        self.ctx.check(lib().pk_ens_resort(self.h))

    def interleave(self):
        """reorder each supercell (FP32) so that consecutive particles sit in different cells"""
        self.ctx.check(lib().pk_ens_interleave(self.h))

    def set_supercells(self, ncps):
        """re-bucket every particle into supercells of ncps cells (pk_ens_set_supercells)"""
        self.ctx.check(lib().pk_ens_set_supercells(self.h, _i3(ncps)))
        self.ncps = tuple(int(v) for v in ncps)

    def save(self, path):
        """raw snapshot of the particles in supercell order (pica_amd.snapshot reads it)"""
        self.ctx.check(lib().pk_ens_save(self.h, os.fsencode(path)))

    def load(self, path):
        """restore a snapshot of an ensemble of the same geometry, species and precision: the
        same particles, bits and order, so a continued run equals the uninterrupted one"""
        self.ctx.check(lib().pk_ens_load(self.h, os.fsencode(path)))

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                lib().pk_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many (ensemble, grid) pairs of one precision and supercell shape stepped together
    (pk_batch_*): one launch per stage for all of them, optionally a captured HIP graph per step.
    The ensembles and grids must outlive the batch."""

    def __init__(self, ctx, pairs):
        self.ctx = ctx
        self.pairs = list(pairs)
        n = len(self.pairs)
        ens = (ctypes.c_void_p * n)(*[e.h.value for e, _ in self.pairs])
        grids = (ctypes.c_void_p * n)(*[g.h.value for _, g in self.pairs])
        h = ctypes.c_void_p()
        ctx.check(lib().pk_batch_create(ctx.h, n, ens, grids, ctypes.byref(h)))
        self.h = h
        ctx._adopt(self)

    def set(self, knob, value):
        self.ctx.check(lib().pk_batch_set(self.h, int(knob), int(value)))

    def get(self, knob):
        v = ctypes.c_int()
        self.ctx.check(lib().pk_batch_get(self.h, int(knob), ctypes.byref(v)))
        return v.value

    def set_shape(self, index, shape):
        """form factor of pair `index`, used by runs with shape=SHAPE_PER_INSTANCE"""
        self.ctx.check(lib().pk_batch_set_shape(self.h, int(index), int(shape)))

    def run(self, dt, nsteps, shape=CIC, dep=DIRECT):
        self.ctx.check(lib().pk_batch_run(self.h, float(dt), shape, dep, int(nsteps)))

    def autotune(self, dt, shape=CIC, dep=DIRECT, trial_steps=3):
        """from the current state (restored after every trial) times supercell sizes, FP32 interleave
        periods and launch settings, keeps the fastest (the ensembles end re-bucketed to its supercell
        size); returns its ms per step"""
        ms = ctypes.c_double()
        self.ctx.check(lib().pk_batch_autotune(self.h, float(dt), shape, dep, int(trial_steps), ctypes.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                lib().pk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Denoiser:
    """The learned current denoiser (pk_dn_*): periodic conv3d 3 -> 16 -> 16 -> 3, BF16 MFMA.
    Host volumes are float32 [3][n0][n1][n2] in network units (scale x J)."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(lib().pk_dn_create(ctx.h, ctypes.byref(h)))
        self.h = h
        ctx._adopt(self)

    @staticmethod
    def num_params():
        return int(lib().pk_dn_num_params())

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params()
        self.ctx.check(lib().pk_dn_set_params(self.h, _p(p)))

    def params(self):
        p = np.empty(self.num_params(), np.float32)
        self.ctx.check(lib().pk_dn_get_params(self.h, _p(p)))
        return p

    def set_scale(self, scale):
        self.ctx.check(lib().pk_dn_set_scale(self.h, float(scale)))

    def init(self, seed):
        """He-normal weights, zero biases"""
        self.ctx.check(lib().pk_dn_init(self.h, int(seed) & 0xFFFFFFFFFFFFFFFF))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty_like(x)
        self.ctx.check(lib().pk_dn_forward_host(self.h, _i3(x.shape[1:]), _p(x), _p(y)))
        return y

    def layer(self, k, x):
        """one convolution layer (1..3) of host input [cin][...]: pre-activation [cout][...]"""
        x = np.ascontiguousarray(x, np.float32)
        cout = 3 if k == 3 else 16
        y = np.empty((cout,) + x.shape[1:], np.float32)
        self.ctx.check(lib().pk_dn_layer_host(self.h, int(k), _i3(x.shape[1:]), _p(x), _p(y)))
        return y

    def grad(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(self.num_params(), np.float32)
        loss = ctypes.c_double()
        self.ctx.check(lib().pk_dn_grad_host(self.h, _i3(x.shape[1:]), _p(x), _p(y), _p(g), ctypes.byref(loss)))
        return loss.value, g

    def grad_batch(self, x, y):
        """mean loss and gradient over x.shape[0] volumes; x, y are [samples][3][n0][n1][n2]"""
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(self.num_params(), np.float32)
        loss = ctypes.c_double()
        self.ctx.check(lib().pk_dn_grad_batch_host(self.h, _i3(x.shape[2:]), int(x.shape[0]), _p(x), _p(y), _p(g),
                                                   ctypes.byref(loss)))
        return loss.value, g

    def add_pair(self, low, high):
        self.ctx.check(lib().pk_dn_add_pair(self.h, low.h, high.h))

    def add_pair_crops(self, low, high, ncrops, edge, seed):
        """ncrops edge^3 crops of the two periodic grids' J as training pairs (the loss leaves out
        voxels within the network's receptive radius of a crop face); returns the offsets [ncrops][3]"""
        off = np.zeros((max(int(ncrops), 1), 3), np.int32)
        self.ctx.check(lib().pk_dn_add_pair_crops(self.h, low.h, high.h, int(ncrops), int(edge),
                                                  int(seed) & 0xFFFFFFFFFFFFFFFF, _p(off)))
        return off[:int(ncrops)]

    def num_pairs(self):
        return int(lib().pk_dn_num_pairs(self.h))

    def clear_pairs(self):
        self.ctx.check(lib().pk_dn_clear_pairs(self.h))

    def train(self, iters, batch, lr, seed=0):
        losses = np.empty(max(int(iters), 1), np.float32)
        self.ctx.check(lib().pk_dn_train(self.h, int(iters), int(batch), float(lr), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                         _p(losses)))
        return losses[:int(iters)]

    def apply(self, grid):
        self.ctx.check(lib().pk_dn_apply(self.h, grid.h))

    def set_charge_gain(self, gain, kmax):
        """radial charge filter of step_denoised: gain g(|k|) on len(gain) points over [0, kmax]
        (radians per cell), gain[0] = 1; an empty gain removes the filter"""
        g = np.ascontiguousarray(gain, np.float64)
        self.ctx.check(lib().pk_dn_set_charge_gain(self.h, int(g.size), float(kmax), _p(g) if g.size else None))

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                lib().pk_dn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rho_save(ens, grid, shape=CIC):
    ens.ctx.check(lib().pk_rho_save(ens.h, grid.h, shape))


def project_continuity(ens, grid, dt, shape=CIC):
    ens.ctx.check(lib().pk_project_continuity(ens.h, grid.h, shape, float(dt)))


def step_denoised(ens, grid, denoiser, dt, shape=CIC, dep=DIRECT):
    ens.ctx.check(lib().pk_step_denoised(ens.h, grid.h, denoiser.h, float(dt), shape, dep))


def particle_step(ens, grid, dt, shape=CIC, dep=DIRECT):
    ens.ctx.check(lib().pk_particle_step(ens.h, grid.h, float(dt), shape, dep))


def step(ens, grid, dt, shape=CIC, dep=DIRECT):
    ens.ctx.check(lib().pk_step(ens.h, grid.h, float(dt), shape, dep))


def step_species(ensembles, grid, dt, shape=CIC, dep=DIRECT):
    """one PIC step of several species (electrons, mobile ions, ...) sharing one grid"""
    hs = (ctypes.c_void_p * len(ensembles))(*[e.h for e in ensembles])
    grid.ctx.check(lib().pk_step_species(hs, len(ensembles), grid.h, float(dt), shape, dep))


def run(ens, grid, dt, nsteps, shape=CIC, dep=DIRECT):
    ens.ctx.check(lib().pk_run(ens.h, grid.h, float(dt), shape, dep, int(nsteps)))


def run_captured(ens, grid, dt, nsteps, shape=CIC, dep=DIRECT):
    """run() with the step captured into a HIP graph and no host synchronisation between steps"""
    ens.ctx.check(lib().pk_run_captured(ens.h, grid.h, float(dt), shape, dep, int(nsteps)))


def deposit_rho(ens, grid, shape=CIC):
    ens.ctx.check(lib().pk_deposit_rho(ens.h, grid.h, shape))


def diag(ens, grid, shape=CIC):
    out = np.zeros(DIAG_N)
    ens.ctx.check(lib().pk_diag(ens.h, grid.h, shape, _p(out)))
    return out


def diag_energies_static(ens, grid):
    """(field energy, kinetic energy) as pk_diag gives them, computed without any device allocation
    (pk_diag_energies_static: a static per-GPU scratch; at most 1,048,576 supercells)"""
    out = np.zeros(2)
    ens.ctx.check(lib().pk_diag_energies_static(ens.h, grid.h, _p(out)))
    return out


def diag_set_reference(ens, grid, shape=CIC):
    ens.ctx.check(lib().pk_diag_set_reference(ens.h, grid.h, shape))


def plasma_parameters(dt, ppc, omega_p_dt=0.1, step=1.0):
    """Electron density n0 = omega_p^2 m / (4 pi e^2) for omega_p * dt = omega_p_dt, and the
    macro-particle weight n0 h^3 / ppc (CGS, pica's constants)."""
    wp = omega_p_dt / dt
    n0 = wp * wp * ELECTRON_MASS / (4.0 * PI * ELECTRON_CHARGE * ELECTRON_CHARGE)
    return n0, n0 * step ** 3 / ppc
