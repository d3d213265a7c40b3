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
batched: the batching form's workload, 64 independent periodic plasmas with cubic edges drawn
from {32,48,64,96,128}, ppc from {4,8,16,32}, drift speeds uniform in [0,0.2]c and CIC or TSC per
plasma (6.3e8 particles, FP32 split positions, Esirkepov), 100 steps, stepped one by one (pk_step
per plasma) and as one batch (pk_batch_run after pk_batch_autotune over supercell size, interleave
period and launch settings); and 512 small plasmas (edges {16,24,32}, ppc {4,8}), 20 steps.
learned: the current denoiser and its charge filter trained on 128^3 two-stream run pairs (4 and
256 ppc of the same initial distribution; --learned-runs pairs of runs, 32 snapshots each, 4 random
32^3 crops per snapshot; 32,000 Adam steps), evaluated at 128^3 on three held-out seeds against a
256-ppc reference: J error (total, longitudinal, transverse) and field-energy error of 4 ppc raw,
denoised and 64 ppc, and the cost of a 64-ppc step over a 4-ppc denoised step.
roofline extras: the FDTD kernels against the HBM roofline (240 B FP64 / 120 B FP32 per
cell-update), and the particle kernel's LDS and global atomic rates against the chip's peaks, from
the per-launch atomic counts of the rocprofv3 summary in profiles/ (commit stamped) and this run's
kernel time.
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


PROFILE = os.path.join("profiles", "particles_tsc_esirkepov_256cube.md")
# FDTD algorithmic bytes per cell-update (B/2, E, B/2): 12 + 9 + 9 values read or written
FDTD_BYTES = {"f64": 240.0, "f32": 120.0}
# peaks of the atomic rates: LDS bandwidth 128 B per clock per CU (a 64-lane FP64 atomic moves
# 512 B), 256 CUs at 2.4 GHz; and the measured conflict-free rate of ds_add_f64 on this part
LDS_F64_ATOMIC_PEAK = 256 * 2.4e9 * 128 / 512          # wave instructions per second
LDS_F64_ATOMIC_MEASURED = 7.1e10


def profile_value(key):
    """a `key: value` line of the profile summary in profiles/ (None if absent); the summary
    names the commit it was measured at in its `commit:` line"""
    try:
        for line in open(os.path.join(ROOT, PROFILE)):
            if line.startswith(key + ":"):
                v = line.split(":", 1)[1].strip()
                if key == "commit":          # a hash such as 24408e4 is not a number
                    return v
                try:
                    return float(v)
                except ValueError:
                    return v
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
    ctx.set_option(P.OPT_INTERLEAVE_F64, args.f64_interleave)
    if args.interleave is not None:
        ctx.set_option(P.OPT_INTERLEAVE, args.interleave)
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
        headline = (args.shape, args.dep, args.n, args.ppc, args.sc) == ("tsc", "esirkepov", 256, 16, 4)
        tpp = profile_value("traffic_bytes_per_particle_%s" % prec) if headline else None
        out["roofline"] = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                           "frac": ach / HBM_PEAK, "traffic": tpp * npart if tpp else None,
                           "traffic_source": "%s at commit %s" % (PROFILE, profile_value("commit")) if tpp else None}
        # atomic rates: per-launch counts from the profile, this run's mean kernel time
        lds = profile_value("lds_atomic_wave_instr_per_particle_%s" % prec) if headline else None
        glb = profile_value("l2_atomic_requests_per_particle_%s" % prec) if headline else None
        if lds or glb:
            out["atomics"] = {"source": "%s at commit %s" % (PROFILE, profile_value("commit"))}
        if lds:
            rate = lds * npart / mean_s
            out["atomics"]["lds_f64_atomic_wave_instr_per_s"] = rate
            out["atomics"]["lds_frac_of_bandwidth_peak"] = rate / LDS_F64_ATOMIC_PEAK
            out["atomics"]["lds_frac_of_measured_conflict_free_rate"] = rate / LDS_F64_ATOMIC_MEASURED
        if glb:
            out["atomics"]["l2_atomic_requests_per_s"] = glb * npart / mean_s
    if "yee_e" in kernels and "yee_b" in kernels:
        # FDTD cell updates per second of the B/2, E, B/2 kernels alone, and their HBM roofline
        fd = (kernels["yee_e"]["ms_mean"] + 2 * kernels["yee_b"]["ms_mean"]) * 1e-3
        out["fdtd_cell_updates_per_s"] = float(np.prod(ncell)) / fd
        a = out["fdtd_cell_updates_per_s"] * FDTD_BYTES[prec]
        out["fdtd_roofline"] = {"bound": "hbm", "achieved": a / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                "frac": a / HBM_PEAK, "bytes_per_cell_update": FDTD_BYTES[prec]}
    ens.close()
    grid.close()
    ctx.close()
    return out


def batch_specs(n=64, seed=3, edges=(32, 48, 64, 96, 128), ppcs=(4, 8, 16, 32)):
    """the batching form's plasmas: per plasma a cubic edge and a ppc drawn uniformly from the
    sets, a drift speed uniform in [0, 0.2]c (as u = v / sqrt(1 - v^2)), CIC or TSC with equal odds
    and its own seed (the draw of tests/helpers.uneven_plasma_specs)"""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        v = 0.2 * rng.random()
        out.append({"edge": int(rng.choice(edges)), "ppc": int(rng.choice(ppcs)), "drift": v / math.sqrt(1.0 - v * v),
                    "shape": int(rng.integers(0, 2)), "seed": 1000 + i})
    return out


def run_batched(args, specs, nsteps):
    """one-by-one vs batched stepping of a batched workload (FP32, Esirkepov, each plasma its own
    form factor and drift), both over the same steps from the same seeded load; aggregate
    particle-steps/s"""
    import pica_amd as P
    ctx = P.Context(0)
    dt = 0.5 / P.C

    def load():
        pairs = []
        for sp in specs:
            n0, w0 = P.plasma_parameters(dt, sp["ppc"])
            g = P.YeeGrid.periodic(ctx, (sp["edge"],) * 3, prec=P.F32)
            e = P.EnsembleSupercells(ctx, (sp["edge"],) * 3, (1.0, 1.0, 1.0), (args.sc,) * 3, prec=P.F32)
            e.load_thermal(sp["ppc"], 0.01, sp["drift"], w0, sp["seed"])
            pairs.append((e, g))
        return pairs

    pairs = load()
    npart = sum(e.size() for e, _ in pairs)
    shapes = [sp["shape"] for sp in specs]
    out = {"instances": len(pairs), "particles": npart, "steps": nsteps,
           "form_factors": {"cic": shapes.count(P.CIC), "tsc": shapes.count(P.TSC)},
           "edges": sorted({sp["edge"] for sp in specs}), "ppc": sorted({sp["ppc"] for sp in specs}),
           "drift_c": [round(min(sp["drift"] / math.sqrt(1 + sp["drift"] ** 2) for sp in specs), 4),
                       round(max(sp["drift"] / math.sqrt(1 + sp["drift"] ** 2) for sp in specs), 4)]}
    for (e, g), sh in zip(pairs, shapes):
        P.step(e, g, dt, sh, P.ESIRKEPOV)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        for (e, g), sh in zip(pairs, shapes):
            P.step(e, g, dt, sh, P.ESIRKEPOV)
    ctx.sync()
    t1 = time.perf_counter()
    out["one_by_one"] = {"value": npart * nsteps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / nsteps}
    print("batched workload one by one: %.3g particle-steps/s" % out["one_by_one"]["value"], file=sys.stderr, flush=True)
    for e, g in pairs:
        e.close()
        g.close()
    pairs = load()
    b = P.Batch(ctx, pairs)
    for i, sh in enumerate(shapes):
        b.set_shape(i, sh)
    tuned_ms = b.autotune(dt, P.SHAPE_PER_INSTANCE, P.ESIRKEPOV, trial_steps=2)
    knobs = {k: b.get(v) for k, v in (("wg_per_cu", P.BATCH_WG_PER_CU), ("order", P.BATCH_ORDER),
                                         ("graph", P.BATCH_GRAPH), ("supercell", P.BATCH_SUPERCELL),
                                         ("interleave", P.BATCH_INTERLEAVE), ("block", P.BATCH_BLOCK))}
    b.run(dt, 1, P.SHAPE_PER_INSTANCE, P.ESIRKEPOV)
    ctx.sync()
    ctx.timer_reset()
    ctx.timer_enable(True)         # only the rebuild count is read: the batched step records no events
    t0 = time.perf_counter()
    b.run(dt, nsteps, P.SHAPE_PER_INSTANCE, P.ESIRKEPOV)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.timer_enable(False)
    out["batched"] = {"value": npart * nsteps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / nsteps,
                      "autotuned_knobs": knobs, "autotune_trial_ms_per_step": tuned_ms,
                      "slack_rebuilds": ctx.timer("rebuild")[1]}
    out["speedup"] = out["batched"]["value"] / out["one_by_one"]["value"]
    print("batched workload as one batch: %.3g particle-steps/s (knobs %s)" % (out["batched"]["value"], knobs),
          file=sys.stderr, flush=True)
    b.close()
    for e, g in pairs:
        e.close()
        g.close()
    ctx.close()
    return out


def run_learned(runs):
    """train the current denoiser on crops of 128^3 run pairs (4 vs 256 ppc) and evaluate it at 128^3"""
    import pica_amd as P
    from pica_amd import learned
    ctx = P.Context(0)
    def log(msg):                      # every 8th training run pair, so that a long leg shows progress
        done = int(msg.split()[2].split("/")[0]) if msg.startswith("training run ") else 0
        if done % 8 == 0:
            print(msg, file=sys.stderr, flush=True)
    report, dn = learned.run(ctx, train_runs=runs, log=log)
    dn.close()
    ctx.close()
    ev = report["eval"]
    m = ev["mean_rel_l2_vs_reference"]
    print("learned: J error vs %d ppc: 4 ppc raw %.4f, denoised in the loop %.4f, 64 ppc %.4f; cost ratio %.2f "
          "(against the FP64 64-ppc step %.2f)"
          % (ev["ppc"]["reference"], m["raw_low"], m["denoised_in_loop"], m["mid"], ev["cost"]["mid_over_low_denoised"],
             ev["cost"]["mid_fp64_over_low_denoised"]), file=sys.stderr, flush=True)
    return report


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()


def cpu_load(n, ppc, sc, seed):
    """the device loader's particles, made by its CPU twin in the oracle (identical inputs)"""
    import oracle as O
    dt = 0.5 / O.C
    wp = 0.1 / dt
    n0 = wp * wp * O.E_MASS / (4.0 * O.PI * O.E_CHARGE * O.E_CHARGE)
    x, u, w, _ = O.load_thermal((n, n, n), (sc,) * 3, (1.0, 1.0, 1.0), ppc, 0.01, 0.1 / math.sqrt(0.99), n0 / ppc,
                                seed, False)
    return x, u, w, dt


def cpu_reference(args):
    """pica's own supercell CIC path (oracle/_ref: pica's headers compiled) on a bounded sample:
    pica has no TSC or Esirkepov, so this is its CIC + direct deposition, FP64."""
    import oracle as O
    if O.ref() is None:
        return None
    n = args.cpu_n
    x, u, w, dt = cpu_load(n, args.ppc, args.sc, 12345)
    threads = cpu_threads()
    t0 = time.perf_counter()
    O.ref_run_cic_direct((n, n, n), (1.0, 1.0, 1.0), (args.sc,) * 3, O.E_CHARGE, O.E_MASS, dt, args.cpu_steps, x, u,
                         w, nthreads=threads)
    t1 = time.perf_counter()
    return {"value": x.shape[0] * args.cpu_steps / (t1 - t0), "unit": "particle-steps/s", "cores": threads,
            "kind": "reference", "cpu_model": cpu_model(),
            "sample": "pica's own FieldInterpolatorCICSupercell + BorisPusher + CurrentDepositorCICSupercell + "
                      "YeeSolver, %d^3 x %d ppc from the seeded loader's CPU twin, %d steps incl. bucketing and "
                      "re-sort glue, FP64 (pica has no TSC/Esirkepov)" % (n, args.ppc, args.cpu_steps)}


def cpu_baseline(args):
    """CPU oracle on a bounded sample of the same workload: sample^3 cells, same ppc, shape,
    depositor and supercells, the device loader's particles (its CPU twin), a few steps, all host
    cores that OMP_NUM_THREADS allows."""
    import oracle as O
    n = args.cpu_n
    x, u, w, dt = cpu_load(n, args.ppc, args.sc, 12345)
    npart = x.shape[0]
    grid = O.Grid.periodic((n, n, n))
    ens = O.Ensemble((n, n, n), (args.sc,) * 3, (1.0, 1.0, 1.0), x, u, w)
    shape = O.TSC if args.shape == "tsc" else O.CIC
    dep = O.ESIRKEPOV if args.dep == "esirkepov" else O.DIRECT
    threads = cpu_threads()
    O.step(grid, ens, dt, shape, dep, threads)     # warm-up
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps):
        O.step(grid, ens, dt, shape, dep, threads)
    t1 = time.perf_counter()
    return {"value": npart * args.cpu_steps / (t1 - t0), "unit": "particle-steps/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": "%d^3 cells x %d ppc (%d particles from the seeded loader's CPU twin, the GPU run's loader), "
                      "%d timed steps of the whole PIC step, FP64, %s + %s, %d^3 supercells"
                      % (n, args.ppc, npart, args.cpu_steps, args.shape.upper(), args.dep, args.sc)}


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
    ap.add_argument("--cpu-n", type=int, default=256, help="CPU baseline grid edge (the workload's own: 256)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-events", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-plasma batched workload")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-denoiser workload")
    ap.add_argument("--learned-runs", type=int, default=48,
                    help="128^3 run pairs the denoiser trains on (the learned form's full set is 200)")
    ap.add_argument("--batch-n", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--interleave", type=int, default=None,
                    help="FP32 interleave period in re-sorts (PK_OPT_INTERLEAVE; default: the library's, 3)")
    ap.add_argument("--f64-interleave", type=int, default=8,
                    help="FP64 reorder period in re-sorts (PK_OPT_INTERLEAVE_F64; 0: off, the library's default)")
    ap.add_argument("--kernel", default="auto", choices=["auto", "per_particle", "cell_sorted"],
                    help="FP32 Esirkepov particle kernel (auto: the per-particle kernel)")
    args = ap.parse_args()
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
        "fdtd_roofline": res.get("fdtd_roofline"),
        "atomics": res.get("atomics"),
        "kernels_ms_mean": {k: v["ms_mean"] for k, v in res["kernels"].items()},
        "kernel_launches": {k: v["launches"] for k, v in res["kernels"].items()},
    }
    if plain is not None:
        line["plain_fp64"] = {"value": plain["particle_steps_per_s"], "ms_per_step": plain["ms_per_step"],
                              "reorder_period": args.f64_interleave,
                              "particle_kernel_rate": plain.get("particle_kernel_rate"),
                              "roofline": plain.get("roofline"),
                              "fdtd_cell_updates_per_s": plain.get("fdtd_cell_updates_per_s"),
                              "fdtd_roofline": plain.get("fdtd_roofline"),
                              "atomics": plain.get("atomics"),
                              "kernels_ms_mean": {k: v["ms_mean"] for k, v in plain["kernels"].items()},
                              "kernel_launches": {k: v["launches"] for k, v in plain["kernels"].items()}}
        line["fp32_speedup_vs_fp64"] = res["particle_steps_per_s"] / plain["particle_steps_per_s"]
        # identical seeded inputs, same number of steps: total-energy difference of the FP32 form
        line["fp32_energy_rel_err_vs_fp64"] = abs(res["energy"] - plain["energy"]) / abs(plain["energy"])
        line["energy_drift_fp64"] = abs(plain["energy"] - plain["energy0"]) / abs(plain["energy0"])
    if not args.no_batched:
        line["batched_workload"] = run_batched(args, batch_specs(args.batch_n), args.batch_steps)
        # many small plasmas, where one launch per stage for all of them matters
        line["batched_small_workload"] = run_batched(args, batch_specs(512, 5, (16, 24, 32), (4, 8)), 20)
    if not args.no_learned:
        line["learned"] = run_learned(args.learned_runs)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args)
        line["gpu_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
        ref = cpu_reference(argparse.Namespace(cpu_n=96, ppc=args.ppc, sc=args.sc, cpu_steps=2))
        if ref:
            line["cpu_baseline_reference"] = ref
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
