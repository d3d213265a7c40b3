"""Learned form driver: training pairs from plain runs, training, and evaluation of the current
denoiser (pk_dn_* in the native library; this module only orchestrates calls into it).

Workload: periodic electron plasmas loaded with a few random long-wavelength velocity modes
(pk_ens_load_modes) on a thermal background, TSC shape, Esirkepov deposition, FP32 split
positions, dt at Courant 0.5, omega_p dt = 0.1.  A training pair is the deposited J of a
low-ppc run and of a high-ppc run of the same modes at the same step; the network maps the first
to the second.  Evaluation on unseen modes compares, against a high-ppc reference run:
  - the raw low-ppc J, the denoised low-ppc J (offline, no feedback) and a mid-ppc run's J;
  - a low-ppc run stepped with pk_step_denoised (denoiser + continuity projection in the loop);
and times a mid-ppc plain step against a low-ppc denoised step.
"""
import math
import time

import numpy as np

from . import (C, ELECTRON_CHARGE, ESIRKEPOV, F32, JX, JY, JZ, TSC, Denoiser, EnsembleSupercells, YeeGrid,
               plasma_parameters, step, step_denoised)

UTH = 0.02        # thermal spread of u = p/(mc)
U_AMP = 0.05      # velocity mode amplitude scale (units of c)
DT = 0.5 / C      # Courant 0.5 at h = 1 cm


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


def plasma(ctx, ncell, ppc, modes, seed):
    _, w = plasma_parameters(DT, ppc)
    g = YeeGrid.periodic(ctx, ncell, prec=F32)
    e = EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), (4, 4, 4), prec=F32)
    e.load_modes(ppc, UTH, 0.0, w, seed, modes)
    return e, g


def current(g):
    """physical J [3][n0][n1][n2], float64"""
    G = 2
    return np.stack([g.get(c)[G:-G, G:-G, G:-G] for c in (JX, JY, JZ)]).astype(np.float64)


def rel_l2(a, ref):
    return float(np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2)))


def generate_pairs(ctx, dn, runs, ncell=(32, 32, 32), ppc_low=4, ppc_high=256, steps=40, every=4, seed=0):
    """runs x (steps / every) training pairs added to dn"""
    rng = np.random.default_rng(seed)
    for r in range(runs):
        modes = random_modes(rng, ncell)
        lo = plasma(ctx, ncell, ppc_low, modes, 100000 * seed + 2 * r)
        hi = plasma(ctx, ncell, ppc_high, modes, 100000 * seed + 2 * r + 1)
        for s in range(1, steps + 1):
            step(*lo, DT, TSC, ESIRKEPOV)
            step(*hi, DT, TSC, ESIRKEPOV)
            if s % every == 0:
                dn.add_pair(lo[1], hi[1])
        for obj in lo + hi:
            obj.close()


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


def evaluate(ctx, dn, ncell=(128, 128, 128), ppc_low=4, ppc_mid=64, ppc_ref=256, steps=32, every=8, seed=99):
    """errors against the ppc_ref run at every `every` steps, and step costs"""
    rng = np.random.default_rng(seed)
    modes = random_modes(rng, ncell)
    scale = network_scale()
    ref = plasma(ctx, ncell, ppc_ref, modes, 7 * seed + 1)
    low = plasma(ctx, ncell, ppc_low, modes, 7 * seed + 2)
    mid = plasma(ctx, ncell, ppc_mid, modes, 7 * seed + 3)
    dnl = plasma(ctx, ncell, ppc_low, modes, 7 * seed + 2)      # same particles as `low`, denoised in the loop
    rows = []
    for s in range(1, steps + 1):
        for e, g in (ref, low, mid):
            step(e, g, DT, TSC, ESIRKEPOV)
        step_denoised(dnl[0], dnl[1], dn, DT, TSC, ESIRKEPOV)
        if s % every == 0:
            jr, jl, jm, jd = current(ref[1]), current(low[1]), current(mid[1]), current(dnl[1])
            jo = dn.forward((jl * scale).astype(np.float32)).astype(np.float64) / scale
            rows.append({"step": s, "raw_low": rel_l2(jl, jr), "denoised_low": rel_l2(jo, jr),
                         "mid": rel_l2(jm, jr), "denoised_in_loop": rel_l2(jd, jr)})
    for obj in ref + low + mid + dnl:
        obj.close()
    # step costs on fresh plasmas
    cost = {}
    mid = plasma(ctx, ncell, ppc_mid, modes, 1)
    cost["mid_plain_ms"] = 1e3 * timed_steps(lambda: step(*mid, DT, TSC, ESIRKEPOV), 10, ctx.sync)
    for obj in mid:
        obj.close()
    low = plasma(ctx, ncell, ppc_low, modes, 2)
    cost["low_plain_ms"] = 1e3 * timed_steps(lambda: step(*low, DT, TSC, ESIRKEPOV), 10, ctx.sync)
    cost["low_denoised_ms"] = 1e3 * timed_steps(lambda: step_denoised(low[0], low[1], dn, DT, TSC, ESIRKEPOV), 10,
                                                ctx.sync)
    # per-kernel HIP-event times of the denoised step
    ctx.timer_reset()
    ctx.timer_enable(True)
    for _ in range(5):
        step_denoised(low[0], low[1], dn, DT, TSC, ESIRKEPOV)
    ctx.sync()
    ctx.timer_enable(False)
    cost["low_denoised_kernels_ms"] = {k: ctx.timer(k)[0] / 5 for k in ("particles", "insert", "denoise", "project",
                                                                        "yee_e", "yee_b", "ghost", "zero")}
    for obj in low:
        obj.close()
    cost["mid_over_low_denoised"] = cost["mid_plain_ms"] / cost["low_denoised_ms"]
    mean = {k: float(np.mean([r[k] for r in rows])) for k in ("raw_low", "denoised_low", "mid", "denoised_in_loop")}
    return {"grid": list(ncell), "ppc": {"low": ppc_low, "mid": ppc_mid, "reference": ppc_ref}, "snapshots": rows,
            "mean_rel_l2_vs_reference": mean, "cost": cost}


def run(ctx, train_runs=48, train_grid=(32, 32, 32), eval_grid=(128, 128, 128), schedule=None, seed=0,
        ppc_target=256):
    """the whole learned-form pipeline; returns (report dict, trained Denoiser)"""
    dn = Denoiser(ctx)
    dn.set_scale(network_scale())
    dn.init(seed + 17)
    t0 = time.perf_counter()
    generate_pairs(ctx, dn, train_runs, train_grid, ppc_high=ppc_target, seed=seed)
    t1 = time.perf_counter()
    losses = train(dn, **({"schedule": schedule} if schedule else {}))
    ctx.sync()
    t2 = time.perf_counter()
    report = {"train": {"grid": list(train_grid), "ppc": {"input": 4, "target": ppc_target}, "pairs": dn.num_pairs(),
                        "iters": int(losses.size),
                        "loss_first": float(np.mean(losses[:20])), "loss_last": float(np.mean(losses[-20:])),
                        "data_s": t1 - t0, "train_s": t2 - t1}}
    report["eval"] = evaluate(ctx, dn, eval_grid)
    dn.clear_pairs()
    return report, dn
