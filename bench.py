#!/usr/bin/env python3
"""Benchmark of the pica PIC step on one MI355X.

Workload (the headline configuration, it fits one GPU): 3D periodic 256^3 grid, dx = 1 cm,
dt at Courant 0.5, 16 electrons per cell (268,435,456 particles) at uniform random positions
with Maxwellian u (sigma 0.01) plus a +-0.1c two-stream drift in x on alternate particles,
omega_p dt = 0.1, TSC form factor, Boris push, Esirkepov (charge-conserving) deposition,
4^3 supercells, re-sorted every step.  A "step" is the whole PIC step: J = 0, fused
interpolation + push + deposition, J fold, B/2-E-B/2 Yee with ghost fills, re-sort (FP32: with
the supercell interleave every 3rd re-sort, the library's default).

value = particles x steps / wall time of the timed steps (inputs resident in HBM).  Defaults: 5
warm-up steps, then the workload's 200 steps timed (the plasma heats as the two-stream
instability grows, so later steps cost more than the first ones).
roofline: the fused particle kernel, algorithmic bytes per particle-step (FP64:
104 + 72/ppc) x particles / its mean launch time from HIP events on the library's stream.
cpu_baseline: the CPU oracle (OpenMP restatement of pica's path) on a bounded sample of the
same workload, on this host's cores.
batched: 64 independent periodic plasmas with edges drawn from {32,48,64,96,128} and ppc from
{4,8,16,32} (4.8e8 particles, FP32 split positions, TSC, Esirkepov), and 512 small ones (edges
{16,24,32}, ppc {4,8}), each stepped one by one (pk_step per plasma) and as one batch
(pk_batch_run after pk_batch_autotune).
learned: the current denoiser trained on 32^3 pairs of 4- and 256-ppc runs (48 runs x 10 snapshots,
32,000 Adam steps), evaluated at 128^3
against a 256-ppc reference: J error of 4 ppc raw, denoised, and 64 ppc, and the cost of a 64-ppc
step over a 4-ppc denoised step.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s (interp+push+deposit) + %HBM roofline, plain vs FP32/batched"
HBM_PEAK = 8.0e12


def alg_bytes(prec, ppc):
    """algorithmic HBM bytes per particle-step of the fused particle kernel"""
    if prec == "f64":
        return 104.0 + 72.0 / ppc     # read x,u,w 56 B + write x,u 48 B; 6 fields read + 3 J written per cell
    return 60.0 + 36.0 / ppc          # read cell,off,u,w 32 B + write cell,off,u 28 B; FP32 grid


def measured_traffic_per_particle(prec):
    """HBM bytes per particle of one particle-kernel launch, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs summarised in profiles/ (calibrated against a plain copy); None if absent"""
    path = os.path.join(ROOT, "profiles", "particles_tsc_esirkepov_256cube.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_particle_%s:" % prec):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def run_gpu(args, prec):
    """load the workload with the seeded device loader, run warmup + timed steps, return rates,
    per-kernel event times and the final energies"""
    import pica_amd as P
    ctx = P.Context(0)
    ctx.set_option(P.OPT_PARTICLE_KERNEL, {"auto": P.KERNEL_AUTO, "per_particle": P.KERNEL_PER_PARTICLE,
                                           "cell_sorted": P.KERNEL_CELL_SORTED}[args.kernel])
    n = args.n
    ncell = (n, n, n)
    ncps = (args.sc, args.sc, args.sc)
    dt = 0.5 / P.C
    n0, w0 = P.plasma_parameters(dt, args.ppc)
    pc = P.F64 if prec == "f64" else P.F32
    grid = P.YeeGrid.periodic(ctx, ncell, prec=pc)
    ens = P.EnsembleSupercells(ctx, ncell, (1.0, 1.0, 1.0), ncps, prec=pc)
    drift = 0.1 / math.sqrt(1.0 - 0.01)
    ens.load_thermal(args.ppc, 0.01, drift, w0, 12345)
    npart = ens.size()
    shape = P.TSC if args.shape == "tsc" else P.CIC
    dep = P.ESIRKEPOV if args.dep == "esirkepov" else P.DIRECT
    d0 = P.diag(ens, grid, shape)
    for _ in range(args.warmup):
        P.step(ens, grid, dt, shape, dep)
    ctx.sync()
    ctx.timer_reset()
    ctx.timer_enable(not args.no_events)
    t0 = time.perf_counter()
    for i in range(args.steps):
        P.step(ens, grid, dt, shape, dep)
        if (i + 1) % 50 == 0:
            print("%s step %d/%d done" % (prec, i + 1, args.steps), file=sys.stderr, flush=True)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.timer_enable(False)
    kernels = {}
    for k in ("particles", "insert", "interleave", "yee_e", "yee_b", "ghost", "zero", "rebuild"):
        ms, cnt = ctx.timer(k)
        if cnt:
            kernels[k] = {"ms_total": ms, "launches": cnt, "ms_mean": ms / cnt}
    d = P.diag(ens, grid, shape)
    out = {
        "prec": prec,
        "particles": npart,
        "wall_s": t1 - t0,
        "ms_per_step": 1e3 * (t1 - t0) / args.steps,
        "particle_steps_per_s": npart * args.steps / (t1 - t0),
        "cell_updates_per_s": float(np.prod(ncell)) * args.steps / (t1 - t0),
        "kernels": kernels,
        "energy0": d0[P.DIAG_FIELD_ENERGY] + d0[P.DIAG_KINETIC_ENERGY],
        "energy": d[P.DIAG_FIELD_ENERGY] + d[P.DIAG_KINETIC_ENERGY],
        "field_energy": d[P.DIAG_FIELD_ENERGY],
        "gauss_max_over_rho": d[P.DIAG_GAUSS_MAX] / max(d[P.DIAG_RHO_MAX], 1e-300),
    }
    if "particles" in kernels:
        mean_s = kernels["particles"]["ms_mean"] * 1e-3
        ach = alg_bytes(prec, args.ppc) * npart / mean_s
        out["particle_kernel_rate"] = npart / mean_s
        tpp = measured_traffic_per_particle(prec) if (args.shape, args.dep, args.n, args.ppc, args.sc) == \
            ("tsc", "esirkepov", 256, 16, 4) else None
        out["roofline"] = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                           "frac": ach / HBM_PEAK, "traffic": tpp * npart if tpp else None}
    if "yee_e" in kernels and "yee_b" in kernels:
        # FDTD cell updates per second of the B/2, E, B/2 kernels alone
        fd = (kernels["yee_e"]["ms_mean"] + 2 * kernels["yee_b"]["ms_mean"]) * 1e-3
        out["fdtd_cell_updates_per_s"] = float(np.prod(ncell)) / fd
    ens.close()
    grid.close()
    ctx.close()
    return out


def batch_specs(n=64, seed=3, edges=(32, 48, 64, 96, 128), ppcs=(4, 8, 16, 32)):
    """a batched workload: (edge, ppc, seed) of each plasma"""
    rng = np.random.default_rng(seed)
    e = rng.choice(list(edges), n)
    p = rng.choice(list(ppcs), n)
    return [(int(a), int(b), 1000 + i) for i, (a, b) in enumerate(zip(e, p))]


def run_batched(args, specs):
    """one-by-one vs batched stepping of a batched workload; aggregate particle-steps/s"""
    import pica_amd as P
    ctx = P.Context(0)
    dt = 0.5 / P.C
    shape = P.TSC if args.shape == "tsc" else P.CIC
    dep = P.ESIRKEPOV if args.dep == "esirkepov" else P.DIRECT
    pairs = []
    for edge, ppc, seed in specs:
        n0, w0 = P.plasma_parameters(dt, ppc)
        g = P.YeeGrid.periodic(ctx, (edge,) * 3, prec=P.F32)
        e = P.EnsembleSupercells(ctx, (edge,) * 3, (1.0, 1.0, 1.0), (args.sc,) * 3, prec=P.F32)
        e.load_thermal(ppc, 0.01, 0.1 / math.sqrt(0.99), w0, seed)
        pairs.append((e, g))
    npart = sum(e.size() for e, _ in pairs)
    out = {"instances": len(pairs), "particles": npart, "steps": args.batch_steps}
    for e, g in pairs:
        P.step(e, g, dt, shape, dep)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        for e, g in pairs:
            P.step(e, g, dt, shape, dep)
    ctx.sync()
    t1 = time.perf_counter()
    out["one_by_one"] = {"value": npart * args.batch_steps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / args.batch_steps}
    print("batched workload one by one: %.3g particle-steps/s" % out["one_by_one"]["value"], file=sys.stderr, flush=True)
    b = P.Batch(ctx, pairs)
    tuned_ms = b.autotune(dt, shape, dep, trial_steps=2)
    knobs = {k: b.get(v) for k, v in (("wg_per_cu", P.BATCH_WG_PER_CU), ("order", P.BATCH_ORDER),
                                         ("graph", P.BATCH_GRAPH))}
    b.run(dt, 1, shape, dep)
    ctx.sync()
    t0 = time.perf_counter()
    b.run(dt, args.batch_steps, shape, dep)
    ctx.sync()
    t1 = time.perf_counter()
    out["batched"] = {"value": npart * args.batch_steps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / args.batch_steps,
                      "autotuned_knobs": knobs, "autotune_trial_ms_per_step": tuned_ms}
    out["speedup"] = out["batched"]["value"] / out["one_by_one"]["value"]
    print("batched workload as one batch: %.3g particle-steps/s" % out["batched"]["value"], file=sys.stderr,
          flush=True)
    b.close()
    for e, g in pairs:
        e.close()
        g.close()
    ctx.close()
    return out


def run_learned():
    """train the current denoiser on 32^3 pairs (4 vs 256 ppc) and evaluate it at 128^3"""
    import pica_amd as P
    from pica_amd import learned
    ctx = P.Context(0)
    report, dn = learned.run(ctx)
    dn.close()
    ctx.close()
    ev = report["eval"]
    print("learned: mean rel L2 vs %d ppc %s, cost ratio %.2f" % (ev["ppc"]["reference"], ev["mean_rel_l2_vs_reference"],
                                                                  ev["cost"]["mid_over_low_denoised"]),
          file=sys.stderr, flush=True)
    return report


def cpu_reference(args):
    """pica's own supercell CIC path (oracle/_ref: pica's headers compiled) on a bounded sample:
    pica has no TSC or Esirkepov, so this is its CIC + direct deposition, FP64."""
    import oracle as O
    if O.ref() is None:
        return None
    rng = np.random.default_rng(8)
    n = args.cpu_n
    ppc = args.ppc
    npart = n ** 3 * ppc
    dt = 0.5 / O.C
    x = rng.random((npart, 3)) * n
    u = rng.standard_normal((npart, 3)) * 0.01
    w = np.full(npart, 1e9)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    t0 = time.perf_counter()
    O.ref_run_cic_direct((n, n, n), (1.0, 1.0, 1.0), (args.sc,) * 3, O.E_CHARGE, O.E_MASS, dt, args.cpu_steps, x, u,
                         w, nthreads=threads)
    t1 = time.perf_counter()
    return {"value": npart * args.cpu_steps / (t1 - t0), "unit": "particle-steps/s", "cores": threads,
            "kind": "reference",
            "sample": "pica's own FieldInterpolatorCICSupercell + BorisPusher + CurrentDepositorCICSupercell + "
                      "YeeSolver, %d^3 x %d ppc, %d steps incl. bucketing and re-sort glue, FP64 (pica has no "
                      "TSC/Esirkepov)" % (n, ppc, args.cpu_steps)}


def cpu_baseline(args):
    """CPU oracle on a bounded sample of the same workload: sample^3 cells, same ppc, shape,
    depositor and supercells, a few steps, all host cores that OMP_NUM_THREADS allows."""
    import oracle as O
    rng = np.random.default_rng(7)
    n = args.cpu_n
    ncell = (n, n, n)
    dt = 0.5 / O.C
    wp = 0.1 / dt
    n0 = wp * wp * O.E_MASS / (4.0 * O.PI * O.E_CHARGE * O.E_CHARGE)
    ppc = args.ppc
    npart = n ** 3 * ppc
    cells = np.repeat(np.arange(n ** 3), ppc)
    x = np.stack([cells // (n * n) + rng.random(npart), (cells // n) % n + rng.random(npart),
                  cells % n + rng.random(npart)], 1).astype(np.float64)
    u = rng.standard_normal((npart, 3)) * 0.01
    u[:, 0] += np.where(np.arange(npart) % 2 == 0, 1.0, -1.0) * 0.1 / math.sqrt(0.99)
    w = np.full(npart, n0 / ppc)
    grid = O.Grid.periodic(ncell)
    ens = O.Ensemble(ncell, (args.sc,) * 3, (1.0, 1.0, 1.0), x, u, w)
    shape = O.TSC if args.shape == "tsc" else O.CIC
    dep = O.ESIRKEPOV if args.dep == "esirkepov" else O.DIRECT
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    O.step(grid, ens, dt, shape, dep, threads)     # warm-up
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps):
        O.step(grid, ens, dt, shape, dep, threads)
    t1 = time.perf_counter()
    return {"value": npart * args.cpu_steps / (t1 - t0), "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": "%d^3 cells x %d ppc (%d particles), %d timed steps of the whole PIC step, FP64, %s + %s, "
                      "%d^3 supercells" % (n, ppc, npart, args.cpu_steps, args.shape.upper(), args.dep, args.sc)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--ppc", type=int, default=16)
    ap.add_argument("--sc", type=int, default=4)
    ap.add_argument("--shape", default="tsc", choices=["cic", "tsc"])
    ap.add_argument("--dep", default="esirkepov", choices=["direct", "esirkepov"])
    ap.add_argument("--prec", default="both", choices=["f64", "f32", "both"],
                    help="f32: the split-position FP32 form (the headline value); f64: the plain FP64 path")
    ap.add_argument("--cpu-n", type=int, default=96)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-events", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-plasma batched workload")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-denoiser workload")
    ap.add_argument("--batch-n", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--kernel", default="auto", choices=["auto", "per_particle", "cell_sorted"],
                    help="FP32 Esirkepov particle kernel (auto: the per-particle kernel)")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="only copy this many GiB device-to-device once (counter calibration for profiles/)")
    args = ap.parse_args()
    if args.copy_gib > 0:
        import torch
        a = torch.ones(int(args.copy_gib * 2 ** 28), device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print("copied %.3f GiB" % args.copy_gib, file=sys.stderr)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (multi-GPU domain decomposition is design-only); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    head = "f64" if args.prec == "f64" else "f32"
    res = run_gpu(args, head)
    print("gpu %s done: %.3g particle-steps/s" % (head, res["particle_steps_per_s"]), file=sys.stderr, flush=True)
    plain = None
    if args.prec == "both":
        plain = run_gpu(args, "f64")
        print("gpu f64 done: %.3g particle-steps/s" % plain["particle_steps_per_s"], file=sys.stderr, flush=True)
    line = {
        "metric": METRIC,
        "value": res["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "f32" else "f64",
        "data": "synthetic (seeded device loader: uniform positions, Maxwellian + two-stream momenta)",
        "config": {"workload": "3D periodic %d^3 Yee grid, %d ppc (%d particles), two-stream +-0.1c, %s, %s, "
                               "%d^3 supercells re-sorted every step, %s; whole PIC step timed"
                               % (args.n, args.ppc, res["particles"], args.shape.upper(), args.dep, args.sc,
                                  "FP32 split positions (int32 cell + FP32 offset), supercells interleaved "
                                  "every 3rd re-sort" if head == "f32" else "FP64"),
                   "grid": [args.n] * 3, "ppc": args.ppc, "particles": res["particles"],
                   "particle_kernel": args.kernel},
        "roofline": res.get("roofline"),
        "particle_kernel_rate": res.get("particle_kernel_rate"),
        "fdtd_cell_updates_per_s": res.get("fdtd_cell_updates_per_s"),
        "kernels_ms_mean": {k: v["ms_mean"] for k, v in res["kernels"].items()},
    }
    if plain is not None:
        line["plain_fp64"] = {"value": plain["particle_steps_per_s"], "ms_per_step": plain["ms_per_step"],
                              "particle_kernel_rate": plain.get("particle_kernel_rate"),
                              "roofline": plain.get("roofline"),
                              "fdtd_cell_updates_per_s": plain.get("fdtd_cell_updates_per_s"),
                              "kernels_ms_mean": {k: v["ms_mean"] for k, v in plain["kernels"].items()}}
        line["fp32_speedup_vs_fp64"] = res["particle_steps_per_s"] / plain["particle_steps_per_s"]
        # identical seeded inputs, same number of steps: total-energy difference of the FP32 form
        line["fp32_energy_rel_err_vs_fp64"] = abs(res["energy"] - plain["energy"]) / abs(plain["energy"])
        line["energy_drift_fp64"] = abs(plain["energy"] - plain["energy0"]) / abs(plain["energy0"])
    if not args.no_batched:
        line["batched_workload"] = run_batched(args, batch_specs(args.batch_n))
        # many small plasmas, where one launch per stage for all of them matters
        line["batched_small_workload"] = run_batched(args, batch_specs(512, 5, (16, 24, 32), (4, 8)))
    if not args.no_learned:
        line["learned"] = run_learned()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args)
        line["gpu_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
        ref = cpu_reference(argparse.Namespace(cpu_n=64, ppc=args.ppc, sc=args.sc, cpu_steps=2))
        if ref:
            line["cpu_baseline_reference"] = ref
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
