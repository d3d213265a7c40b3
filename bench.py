#!/usr/bin/env python3
"""Benchmark of the pica PIC step on one MI355X.

Workload (the headline configuration, it fits one GPU): 3D periodic 256^3 grid, dx = 1 cm,
dt at Courant 0.5, 16 electrons per cell (268,435,456 particles) at uniform random positions
with Maxwellian u (sigma 0.01) plus a +-0.1c two-stream drift in x on alternate particles,
omega_p dt = 0.1, TSC form factor, Boris push, Esirkepov (charge-conserving) deposition,
4^3 supercells, re-sorted every step.  A "step" is the whole PIC step: J = 0, fused
interpolation + push + deposition, J fold, B/2-E-B/2 Yee with ghost fills, re-sort.

value = particles x steps / wall time of the timed steps (inputs resident in HBM).
roofline: the fused particle kernel, algorithmic bytes per particle-step (FP64:
104 + 72/ppc) x particles / its mean launch time from HIP events on the library's stream.
cpu_baseline: the CPU oracle (OpenMP restatement of pica's path) on a bounded sample of the
same workload, on this host's cores.
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


def run_gpu(args, prec):
    import pica_amd as P
    ctx = P.Context(0)
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
    for _ in range(args.warmup):
        P.step(ens, grid, dt, shape, dep)
    ctx.sync()
    ctx.timer_reset()
    ctx.timer_enable(not args.no_events)
    t0 = time.perf_counter()
    for i in range(args.steps):
        P.step(ens, grid, dt, shape, dep)
        print("step %d/%d done" % (i + 1, args.steps), file=sys.stderr, flush=True) if (i + 1) % 20 == 0 else None
    ctx.sync()
    t1 = time.perf_counter()
    ctx.timer_enable(False)
    kernels = {}
    for k in ("particles", "insert", "yee_e", "yee_b", "ghost", "zero", "rebuild"):
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
        "energy": {"field": d[P.DIAG_FIELD_ENERGY], "kinetic": d[P.DIAG_KINETIC_ENERGY]},
    }
    if "particles" in kernels:
        mean_s = kernels["particles"]["ms_mean"] * 1e-3
        ach = alg_bytes(prec, args.ppc) * npart / mean_s
        out["particle_kernel_rate"] = npart / mean_s
        out["roofline"] = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                           "frac": ach / HBM_PEAK, "traffic": None}
    ens.close()
    grid.close()
    ctx.close()
    return out


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
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--ppc", type=int, default=16)
    ap.add_argument("--sc", type=int, default=4)
    ap.add_argument("--shape", default="tsc", choices=["cic", "tsc"])
    ap.add_argument("--dep", default="esirkepov", choices=["direct", "esirkepov"])
    ap.add_argument("--prec", default="f64", choices=["f64", "f32"])
    ap.add_argument("--cpu-n", type=int, default=96)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-events", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (multi-GPU domain decomposition is design-only); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    res = run_gpu(args, args.prec)
    print("gpu done: %.3g particle-steps/s" % res["particle_steps_per_s"], file=sys.stderr, flush=True)
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
        "dtype": "f64" if args.prec == "f64" else "f32",
        "data": "synthetic (seeded device loader: uniform positions, Maxwellian + two-stream momenta)",
        "config": {"workload": "3D periodic %d^3 Yee grid, %d ppc (%d particles), two-stream +-0.1c, %s, %s, "
                               "%d^3 supercells re-sorted every step, %s; whole PIC step timed"
                               % (args.n, args.ppc, res["particles"], args.shape.upper(), args.dep, args.sc,
                                  args.prec.upper()),
                   "grid": [args.n] * 3, "ppc": args.ppc, "particles": res["particles"]},
        "roofline": res.get("roofline"),
        "particle_kernel_rate": res.get("particle_kernel_rate"),
        "cell_updates_per_s": res["cell_updates_per_s"],
        "kernels_ms_mean": {k: v["ms_mean"] for k, v in res["kernels"].items()},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args)
        line["gpu_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
