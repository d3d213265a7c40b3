"""Learned form driver: training pairs from plain runs, training, and evaluation of the current
denoiser (pk_dn_* in the native library; this module only orchestrates calls into it).

Workload: 128^3 periodic two-stream electron plasmas (Maxwellian u with sigma 0.01, drift +-0.1c
in x on alternate particles, as the plain 256^3 workload) carrying a few seeded long-wavelength
velocity modes (pk_ens_load_modes), TSC shape, Esirkepov deposition, FP32 split positions, dt at
Courant 0.5, omega_p dt = 0.1.  A run pair is a 4-ppc and a 256-ppc plasma of the same initial
distribution (the same modes, each its own particles); at each of 32 snapshots (every step) a few
random 32^3 crops of both currents become training pairs, and the network maps the first to the
second.  The modes give the pair a common signal: without them the two runs' currents are
independent noise.  A crop is not periodic, so the loss leaves out its voxels within the
network's receptive radius (3) of a face.  The charge filter: the projection of the denoised step fixes the longitudinal part of J by
continuity, so without a filter that part keeps the low-ppc charge noise whatever the network
does.  A radial filter S with gain g(|k|) is learned from the training pairs as the Wiener gain of
the longitudinal current per |k| shell, g = sum Re(DJ_high conj DJ_low) / sum |DJ_low|^2 (DJ the
Fourier transform of the divergence, weighted by 1 / |lambda|, lambda the Laplacian's eigenvalue),
clipped to [0, 1], g(0) = 1; the denoised step then keeps continuity, and Gauss's law, for the
filtered charge S rho exactly.

Evaluation on unseen modes and held-out seeds compares, against a high-ppc reference run, the J
error (total, and its longitudinal and transverse parts) and the field energy of:
  - the raw low-ppc run, the denoised low-ppc J (offline, no feedback) and a mid-ppc run;
  - a low-ppc run stepped with pk_step_denoised (denoiser + filtered-continuity projection);
and times a mid-ppc plain step against a low-ppc denoised step.
"""
import math
import time

import numpy as np

from . import (C, DIAG_FIELD_ENERGY, ELECTRON_CHARGE, ESIRKEPOV, F32, F64, JX, JY, JZ, OPT_INTERLEAVE, TSC,
               Denoiser, EnsembleSupercells, YeeGrid, diag, plasma_parameters, step, step_denoised)

UTH = 0.01        # thermal spread of u = p/(mc)
U_DRIFT = 0.1 / math.sqrt(1.0 - 0.1 * 0.1)    # u of the +-0.1c streams
U_AMP = 0.05      # velocity mode amplitude scale (units of c)
DT = 0.5 / C      # Courant 0.5 at h = 1 cm
# FP32 interleave period of the timed low-ppc steps: at 128^3 x 4 ppc a 4^3 supercell fills two
# wavefronts and its order decays slowly, so the particle stage plus its re-sort takes 0.82 ms per
# step at periods 6-12 against 0.85 ms at the library's default 3, which stays the best at 64 ppc
# (7.30 ms against 7.38 at 4 and 7.52 at 6)
LOW_INTERLEAVE = 8


def random_modes(rng, ncell, nmodes=3, amp=U_AMP):
    """rows (kx, ky, kz, ax, ay, az, phase); wavelengths of at least 16 cells"""
    kmax = [max(1, n // 16) for n in ncell]
    rows = []
    for _ in range(nmodes):
        while True:
            k = np.array([rng.integers(-m, m + 1) for m in kmax])
            if np.any(k != 0):
                break
        a = rng.standard_normal(3)
        a *= amp * rng.uniform(0.5, 1.0) / np.linalg.norm(a)
        rows.append([k[0], k[1], k[2], a[0], a[1], a[2], rng.uniform(0.0, 2.0 * math.pi)])
    return np.array(rows, np.float64)


def network_scale():
    """network units = J / (|e| n0 c U_AMP), so a mode's current is O(1)"""
    n0, _ = plasma_parameters(DT, 1)
    return 1.0 / (abs(ELECTRON_CHARGE) * n0 * C * U_AMP)


def plasma(ctx, ncell, ppc, modes, seed, ncps=(4, 4, 4), prec=F32):
    _, w = plasma_parameters(DT, ppc)
    g = YeeGrid.periodic(ctx, ncell, prec=prec)
    e = EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps, prec=prec)
    e.load_modes(ppc, UTH, U_DRIFT, w, seed, modes)
    return e, g


def current(g):
    """physical J [3][n0][n1][n2], float64"""
    G = 2
    return np.stack([g.get(c)[G:-G, G:-G, G:-G] for c in (JX, JY, JZ)]).astype(np.float64)


def rel_l2(a, ref):
    return float(np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2)))


def _ops(n):
    """Fourier symbols of D (forward differences), G = -D^T and L = D G on a periodic n-grid, unit steps"""
    K = np.meshgrid(*[2.0 * np.pi * np.fft.fftfreq(m) for m in n], indexing="ij")
    Dh = [np.exp(1j * k) - 1.0 for k in K]
    Gh = [1.0 - np.exp(-1j * k) for k in K]
    lam = sum(d * g for d, g in zip(Dh, Gh)).real
    return K, Dh, Gh, lam


def div_hat(j):
    """Fourier transform of D J (the divergence at the cell-centre nodes) of a physical J [3][n]"""
    _, Dh, _, _ = _ops(j.shape[1:])
    return sum(Dh[c] * np.fft.fftn(j[c]) for c in range(3))


def split_lt(j):
    """longitudinal and transverse parts of J: J_L = G L^-1 D J, J_T = J - J_L (the split under
    which the projection acts: it leaves J_T and sets D J)"""
    _, Dh, Gh, lam = _ops(j.shape[1:])
    lam = lam.copy()
    lam.flat[0] = 1.0
    phi = sum(Dh[c] * np.fft.fftn(j[c]) for c in range(3)) / lam
    phi.flat[0] = 0.0
    jl = np.stack([np.fft.ifftn(Gh[c] * phi).real for c in range(3)])
    return jl, j - jl


class ChargeGainFit:
    """accumulates, over training pairs, the per-|k|-shell sums of the charge filter's Wiener gain:
    A = sum |DJ_low|^2 / |lambda|, B = sum Re(DJ_high conj DJ_low) / |lambda|"""

    def __init__(self, nbins=33):
        self.nbins = nbins
        self.kmax = float(np.pi * np.sqrt(3.0))
        self.A = np.zeros(nbins)
        self.B = np.zeros(nbins)

    def add(self, j_low, j_high):
        K, _, _, lam = _ops(j_low.shape[1:])
        kabs = np.sqrt(sum(k * k for k in K))
        nz = lam != 0
        w = np.zeros_like(lam)
        w[nz] = 1.0 / np.abs(lam[nz])
        dl, dh = div_hat(j_low), div_hat(j_high)
        b = np.clip(np.rint(kabs / self.kmax * (self.nbins - 1)).astype(int), 0, self.nbins - 1)
        self.A += np.bincount(b[nz], (np.abs(dl) ** 2 * w)[nz], self.nbins)
        self.B += np.bincount(b[nz], ((dh * np.conj(dl)).real * w)[nz], self.nbins)

    def fit(self):
        """(gain, kmax): gain on nbins points over [0, kmax], g(0) = 1, shells without data
        interpolated from their neighbours"""
        have = self.A > 0
        g = np.where(have, self.B / np.where(have, self.A, 1.0), np.nan)
        g = np.clip(g, 0.0, 1.0)
        g[0] = 1.0
        idx = np.arange(self.nbins)
        ok = ~np.isnan(g)
        g = np.interp(idx, idx[ok], g[ok])
        return g, self.kmax


def generate_pairs(ctx, dn, runs, ncell=(128, 128, 128), ppc_low=4, ppc_high=256, snapshots=32, every=1, crops=4,
                   edge=32, seed=0, cfit=None, cfit_runs=4, cfit_every=8, log=None):
    """runs x snapshots x crops training pairs added to dn: `runs` seeded run pairs, a snapshot every
    `every` steps, `crops` random edge^3 crops of both currents per snapshot (the whole periodic grid
    when edge covers it). The charge-filter fit cfit takes the whole currents of every cfit_every-th
    snapshot of the first cfit_runs runs."""
    rng = np.random.default_rng(seed)
    whole = all(edge >= n for n in ncell)
    for r in range(runs):
        modes = random_modes(rng, ncell)
        lo = plasma(ctx, ncell, ppc_low, modes, 100000 * seed + 2 * r)
        hi = plasma(ctx, ncell, ppc_high, modes, 100000 * seed + 2 * r + 1)
        for s in range(1, snapshots * every + 1):
            step(*lo, DT, TSC, ESIRKEPOV)
            step(*hi, DT, TSC, ESIRKEPOV)
            if s % every:
                continue
            if whole:
                dn.add_pair(lo[1], hi[1])
            else:
                dn.add_pair_crops(lo[1], hi[1], crops, edge, (seed << 32) + (r << 12) + s)
            if cfit is not None and r < cfit_runs and (s // every) % cfit_every == 0:
                cfit.add(current(lo[1]), current(hi[1]))
        for obj in lo + hi:
            obj.close()
        if log:
            log("training run %d/%d done, %d pairs" % (r + 1, runs, dn.num_pairs()))


def train(dn, schedule=((16000, 2e-3), (12000, 5e-4), (4000, 1e-4)), batch=8, seed=1):
    """Adam over the pairs; returns the per-step losses. 32,000 steps: at 8,000 the denoised 4-ppc error
    at 128^3 is 0.080, at 16,000 0.077, at 32,000 0.074 (the 64-ppc error is 0.076); batch 16 at
    16,000 steps (the same work) gives 0.076, so the steps, not the samples, matter."""
    losses = []
    for k, (iters, lr) in enumerate(schedule):
        losses.append(dn.train(iters, batch, lr, seed + k))
    return np.concatenate(losses)


def timed_steps(fn, nsteps, sync):
    fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        fn()
    sync()
    return (time.perf_counter() - t0) / nsteps


def evaluate(ctx, dn, ncell=(128, 128, 128), ppc_low=4, ppc_mid=64, ppc_ref=256, steps=32, every=8, seed=99,
             low_ncps=(4, 4, 4)):
    """errors against the ppc_ref run at every `every` steps (J total / longitudinal / transverse, and
    the field energy), for one held-out seed"""
    rng = np.random.default_rng(seed)
    modes = random_modes(rng, ncell)
    scale = network_scale()
    ref = plasma(ctx, ncell, ppc_ref, modes, 7 * seed + 1)
    low = plasma(ctx, ncell, ppc_low, modes, 7 * seed + 2, low_ncps)
    mid = plasma(ctx, ncell, ppc_mid, modes, 7 * seed + 3)
    dnl = plasma(ctx, ncell, ppc_low, modes, 7 * seed + 2, low_ncps)  # same particles as `low`, denoised in the loop
    rows = []
    for s in range(1, steps + 1):
        for e, g in (ref, low, mid):
            step(e, g, DT, TSC, ESIRKEPOV)
        step_denoised(dnl[0], dnl[1], dn, DT, TSC, ESIRKEPOV)
        if s % every == 0:
            jr, jl, jm, jd = current(ref[1]), current(low[1]), current(mid[1]), current(dnl[1])
            jo = dn.forward((jl * scale).astype(np.float32)).astype(np.float64) / scale
            rl, rt = split_lt(jr)
            row = {"step": s}
            for name, j in (("raw_low", jl), ("denoised_low", jo), ("mid", jm), ("denoised_in_loop", jd)):
                a, b = split_lt(j)
                row[name] = rel_l2(j, jr)
                row[name + "_L"] = rel_l2(a, rl)
                row[name + "_T"] = rel_l2(b, rt)
            wref = diag(ref[0], ref[1], TSC)[DIAG_FIELD_ENERGY]
            for name, (e, g) in (("raw_low", low), ("mid", mid), ("denoised_in_loop", dnl)):
                row[name + "_field_energy"] = abs(diag(e, g, TSC)[DIAG_FIELD_ENERGY] - wref) / wref
            rows.append(row)
    for obj in ref + low + mid + dnl:
        obj.close()
    return rows


def costs(ctx, dn, ncell=(128, 128, 128), ppc_low=4, ppc_mid=64, low_ncps=(4, 4, 4), seed=5):
    """ms per step of a mid-ppc plain step (FP32 split, as the denoised step, and FP64, pica's
    precision) and of a low-ppc denoised step (fresh plasmas), with the per-kernel HIP-event times of
    the denoised step over one interleave period; the low-ppc legs run at LOW_INTERLEAVE, the
    mid-ppc one at the context's period"""
    modes = random_modes(np.random.default_rng(seed), ncell)
    cost = {}
    mid = plasma(ctx, ncell, ppc_mid, modes, 1)
    cost["mid_plain_ms"] = 1e3 * timed_steps(lambda: step(*mid, DT, TSC, ESIRKEPOV), 10, ctx.sync)
    for obj in mid:
        obj.close()
    mid = plasma(ctx, ncell, ppc_mid, modes, 1, prec=F64)
    cost["mid_plain_fp64_ms"] = 1e3 * timed_steps(lambda: step(*mid, DT, TSC, ESIRKEPOV), 10, ctx.sync)
    for obj in mid:
        obj.close()
    period = ctx.get_option(OPT_INTERLEAVE)
    ctx.set_option(OPT_INTERLEAVE, LOW_INTERLEAVE)
    try:
        low = plasma(ctx, ncell, ppc_low, modes, 2, low_ncps)
        cost["low_plain_ms"] = 1e3 * timed_steps(lambda: step(*low, DT, TSC, ESIRKEPOV), 10, ctx.sync)
        cost["low_denoised_ms"] = 1e3 * timed_steps(
            lambda: step_denoised(low[0], low[1], dn, DT, TSC, ESIRKEPOV), 10, ctx.sync)
        ctx.timer_reset()
        ctx.timer_enable(True)
        for _ in range(LOW_INTERLEAVE):
            step_denoised(low[0], low[1], dn, DT, TSC, ESIRKEPOV)
        ctx.sync()
        ctx.timer_enable(False)
    finally:
        ctx.set_option(OPT_INTERLEAVE, period)
    cost["interleave_period"] = {"low": LOW_INTERLEAVE, "mid": period}
    names = ("particles", "insert", "denoise", "project", "yee_e", "yee_b", "ghost", "zero", "interleave")
    cost["low_denoised_kernels_ms"] = {k: ctx.timer(k)[0] / LOW_INTERLEAVE for k in names}
    for obj in low:
        obj.close()
    cost["mid_over_low_denoised"] = cost["mid_plain_ms"] / cost["low_denoised_ms"]
    cost["mid_fp64_over_low_denoised"] = cost["mid_plain_fp64_ms"] / cost["low_denoised_ms"]
    cost["low_supercells"] = list(low_ncps)
    return cost


def run(ctx, train_runs=48, train_grid=(128, 128, 128), eval_grid=(128, 128, 128), schedule=None, seed=0,
        ppc_target=256, eval_seeds=(99, 199, 299), snapshots=32, crops=4, edge=32, log=None):
    """the whole learned-form pipeline; returns (report dict, trained Denoiser)"""
    dn = Denoiser(ctx)
    dn.set_scale(network_scale())
    dn.init(seed + 17)
    cfit = ChargeGainFit()
    t0 = time.perf_counter()
    generate_pairs(ctx, dn, train_runs, train_grid, ppc_high=ppc_target, snapshots=snapshots, crops=crops, edge=edge,
                   seed=seed, cfit=cfit, log=log)
    t1 = time.perf_counter()
    losses = train(dn, **({"schedule": schedule} if schedule else {}))
    gain, kmax = cfit.fit()
    dn.set_charge_gain(gain, kmax)
    ctx.sync()
    t2 = time.perf_counter()
    report = {"train": {"grid": list(train_grid), "ppc": {"input": 4, "target": ppc_target}, "runs": train_runs,
                        "snapshots_per_run": snapshots, "crops_per_snapshot": crops, "crop_edge": edge,
                        "drift_c": 0.1, "uth": UTH, "pairs": dn.num_pairs(),
                        "iters": int(losses.size),
                        "loss_first": float(np.mean(losses[:20])), "loss_last": float(np.mean(losses[-20:])),
                        "charge_gain": {"kmax": kmax, "gain": [round(float(v), 4) for v in gain]},
                        "data_s": t1 - t0, "train_s": t2 - t1}}
    dn.clear_pairs()
    rows = []
    for sd in eval_seeds:
        rows += [dict(r, seed=sd) for r in evaluate(ctx, dn, eval_grid, seed=sd)]
    keys = [k for k in rows[0] if k not in ("step", "seed")]
    report["eval"] = {"grid": list(eval_grid), "ppc": {"low": 4, "mid": 64, "reference": 256},
                      "held_out_seeds": list(eval_seeds), "snapshots": rows,
                      "mean_rel_l2_vs_reference": {k: float(np.mean([r[k] for r in rows])) for k in keys},
                      "cost": costs(ctx, dn, eval_grid)}
    return report, dn


def main(argv=None):
    """python -m pica_amd.learned [--runs N]: the whole pipeline on device 0, one JSON report line"""
    import argparse
    import json

    from . import Context
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--runs", type=int, default=200, help="128^3 training run pairs (4 and 256 ppc)")
    ap.add_argument("--snapshots", type=int, default=32)
    ap.add_argument("--crops", type=int, default=4)
    args = ap.parse_args(argv)
    ctx = Context(0)
    report, dn = run(ctx, train_runs=args.runs, snapshots=args.snapshots, crops=args.crops,
                     log=lambda m: print(m, flush=True))
    dn.close()
    ctx.close()
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
