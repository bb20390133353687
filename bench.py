"""Time-steps/s of the semi-implicit dynamical-core step on one MI355X.

Workload: the pseudo-incompressible mountain wave at 256 x 256 x 128 cells
(isothermal 250 K, lx = ly = 200 km, lz = 30 km with sinh stretching, one
Gaussian hill 500 m high and 10 km wide, u0 = 10 m/s, f = 1e-4 s^-1, sponge over
the top third, BiCGSTAB with the reference preconditioner to 1e-8), FP64.
A step is one pass of the time loop of integrate (two Poisson solves).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|mixed]
                       [--no-cpu-baseline] [--no-fp64-check]
The mixed-precision configuration (16 seeded hills at 512 x 512 x 128, error
against FP64 after 100 steps):
  python bench.py --workload gaussian_hills --nx 512 --precision mixed --preconditioner multigrid
                  --storage bf16 --mg-coarse-steps 2 --initial-guess previous --steps 100 --no-cpu-baseline
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from __graft_entry__ import load_package  # noqa: E402

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the dot-fused operator apply (k_stencil_pair<0, ndot>)
# measured by rocprofv3 FETCH_SIZE and WRITE_SIZE passes (scaled by the
# counters' readings for a plain copy of known size: x2.0 and x1.0), keyed by
# (nx, ny, nz, precision); summary and commands in
# profiles/k_stencil_pair_mountain_wave_256x256x128_fp64_hbm_traffic.md.
APPLY_TRAFFIC_BYTES = {(256, 256, 128, "fp64"): 1.930e9}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="mountain_wave", choices=["mountain_wave", "gaussian_hills", "ensemble"],
                   help="mountain_wave: one Gaussian hill (the metric's workload); gaussian_hills: the same "
                        "physics over 16 seeded hills (the mixed-precision configuration, run at --nx 512); "
                        "ensemble: --members seeded bubble members sharing the GPU, against one at a time")
    p.add_argument("--members", type=int, default=16, help="ensemble workload: number of members")
    p.add_argument("--member-x", type=int, default=0,
                   help="ensemble workload: x_size = y_size of every member (default: the member's own)")
    p.add_argument("--member-z", type=int, default=0,
                   help="ensemble workload: z_size of every member (default: the member's own)")
    p.add_argument("--ensemble-modes", default="alone,streams,batched",
                   help="ensemble workload: comma-separated runs among alone (one member at a time), streams "
                        "(all members at once, each solve on its own) and batched (all at once, solves batched)")
    p.add_argument("--model", default="PseudoIncompressible", choices=["PseudoIncompressible", "Compressible"],
                   help="dynamic equations of the mountain-wave / hill workloads")
    p.add_argument("--nx", type=int, default=256)
    p.add_argument("--nz", type=int, default=128)
    p.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    p.add_argument("--flux-precision", default=None, choices=["fp64", "mixed"],
                   help="arithmetic of the RK stages' fluxes (default: fp64 with --precision fp64, mixed (FP32) "
                        "with --precision mixed)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--cpu-sample-nx", type=int, default=128)
    p.add_argument("--pc-iterations", type=int, default=2,
                   help="pseudo-time steps of the preconditioner (the reference's default is 2)")
    p.add_argument("--dtau", type=float, default=1.0, help="pseudo-time step of the preconditioner (default 1)")
    p.add_argument("--preconditioner", default="line_relaxation", choices=["line_relaxation", "multigrid"],
                   help="mixed precision: the reference's line relaxation or the multigrid V-cycle")
    p.add_argument("--initial-guess", default="zero", choices=["zero", "previous", "extrapolated", "learned"],
                   help="mixed precision: start each solve from zero, from the same stage's previous solution, or "
                        "from the learned network's guess (trained first, on held-out terrain)")
    p.add_argument("--mg-smoothing-steps", type=int, default=2, help="multigrid pre- and post-smoothing steps")
    p.add_argument("--mg-coarse-steps", type=int, default=20, help="multigrid steps on the coarsest level")
    p.add_argument("--storage", default="fp32", choices=["fp32", "bf16"],
                   help="mixed precision: storage of the finest level's coefficients in the preconditioner")
    p.add_argument("--train-runs", type=int, default=3, help="learned networks: training runs (random hills)")
    p.add_argument("--train-steps", type=int, default=12, help="learned networks: time steps per training run")
    p.add_argument("--train-epochs", type=int, default=10, help="learned guess: passes over the samples (error loss)")
    p.add_argument("--train-loss", default="residual", choices=["residual", "error"],
                   help="learned guess: residual |A x0 - b|^2 (online, --train-inner Adam steps per solve) or "
                        "the squared error of the solution change")
    p.add_argument("--train-inner", type=int, default=4, help="learned networks: Adam steps per solve")
    p.add_argument("--train-objective", default="energy", choices=["residual", "energy"],
                   help="learned networks: the residual |A x - v|^2 or the energy norm of the error under -A")
    p.add_argument("--stage-iterations", type=int, default=1,
                   help="learned stage: leading BiCGSTAB iterations whose preconditioner applications it follows")
    p.add_argument("--train-lr", type=float, default=1e-4,
                   help="learned networks: Adam learning rate")
    p.add_argument("--no-variants", action="store_true",
                   help="plain FP64 run: skip the mixed-precision, multigrid and learned lines")
    p.add_argument("--no-fp64-check", action="store_true",
                   help="mixed precision: skip the FP64 rerun that measures the field differences")
    return p.parse_args()


def workload(pcf, args, iterations, precision):
    """Namelists of the benchmarked workload at --nx x --nx x --nz."""
    if args.workload == "gaussian_hills":
        nl = pcf.examples.gaussian_hills_3d(x_size=args.nx, y_size=args.nx, z_size=args.nz, iterations=iterations,
                                            precision=precision)
    else:
        nl = pcf.examples.mountain_wave_3d(x_size=args.nx, y_size=args.nx, z_size=args.nz, iterations=iterations,
                                           precision=precision)
    nl.atmosphere.model = args.model
    return nl


def run_method(pcf, args, precision, preconditioner, guess, net=None, keep=False, storage=None, mg_coarse=None,
               relative=False, stage_net=None, fluxes="fp64", profile_apply=True):
    """Setup, initial cleaning, warmup, then the timed steps of one method,
    with HIP events around the launches of the operator apply only (its
    average duration gives the roofline figure; events around every launch
    would delay the launch-heavy preconditioners), then one more step with
    events around every kernel group for the breakdown."""
    nl = workload(pcf, args, args.steps + args.warmup + 1, precision)
    nl.poisson.initial_guess = guess
    nl.poisson.preconditioner_type = preconditioner
    nl.discretization.flux_precision = fluxes
    nl.poisson.tolerance_is_relative = relative
    nl.poisson.preconditioner_iterations = args.pc_iterations
    nl.poisson.dtau = args.dtau
    if precision == "mixed":
        nl.poisson.mg_smoothing_steps = args.mg_smoothing_steps
        nl.poisson.mg_coarse_steps = mg_coarse or args.mg_coarse_steps
        nl.poisson.preconditioner_storage = storage or args.storage
    tag = (f"{precision}/{preconditioner}/{nl.poisson.preconditioner_storage}/{guess}" + ("+stage" if stage_net else "")
           + ("/fp32-fluxes" if fluxes == "mixed" else ""))
    t0 = time.perf_counter()
    st = pcf.State(nl)
    st.solver_graphs(True)  # the bench steps one state at a time from this thread
    if net is not None:
        pcf.learned.attach(st, net)
    if stage_net is not None:
        pcf.learned.attach_stage(st, stage_net, args.stage_iterations)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    n_clean = st.initial_cleaning()
    t_clean = time.perf_counter() - t0
    print(f"# [{tag}] setup {t_setup:.2f}s, initial cleaning {t_clean:.2f}s ({n_clean} iterations)",
          file=sys.stderr, flush=True)
    for w in range(args.warmup):
        info = st.step()
        print(f"# [{tag}] warmup step {w + 1}: dt {info.dt:.3f}s iterations {list(info.niter)}", file=sys.stderr,
              flush=True)
    if profile_apply:
        # events around the operator apply only (the roofline kernel); a
        # profiled state launches its BiCGSTAB iterations kernel by kernel
        st.profile(True, groups=["apply_total_dot"])
    iters = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        info = st.step()
        iters.append(list(info.niter))
        print(f"# [{tag}] step {s + 1}: dt {info.dt:.3f}s iterations {list(info.niter)}", file=sys.stderr,
              flush=True)
    elapsed = time.perf_counter() - t0
    prof = st.profile_report()
    st.profile(True)
    info = st.step()
    breakdown = st.profile_report()
    st.profile(False)
    print(f"# [{tag}] breakdown step: iterations {list(info.niter)}", file=sys.stderr, flush=True)
    res = {"nl": nl, "iters": iters, "prof": prof, "breakdown": breakdown, "elapsed": elapsed,
           "stage_tuning": st.stage_tuning(),
           "setup_s": round(t_setup, 3), "initial_cleaning_s": round(t_clean, 3),
           "value": args.steps / elapsed, "ms_per_step": elapsed / args.steps * 1e3,
           "iterations_per_solve": sum(sum(i) for i in iters) / (2 * len(iters))}
    if keep:
        res["state"] = st
    else:
        st.close()
    return res


def field_difference(st_a, st_b):
    """Max-norm differences relative to the field's max norm (the Exner
    pressure up to a constant: the null space of the pressure operator)."""
    out = {}
    for f in ("RHO", "RHOP", "U", "V", "W", "PIP"):
        a, b = st_a.get(f), st_b.get(f)
        d = a - b
        if f == "PIP":
            d = d - d[3:-3, 3:-3, 3:-3].mean()
        scale = float(abs(b).max()) or 1.0
        out[f] = float(abs(d).max()) / scale
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported "
              "(the domain is not decomposed across GPUs in this build)", file=sys.stderr)
        sys.exit(2)
    pcf = load_package()
    if args.workload == "ensemble":
        ensemble_bench(pcf, args)
        return
    plain = args.precision == "fp64"
    training = None
    net = None
    if args.initial_guess == "learned":
        training = train_guess(pcf, args)
        net = training.pop("net", None) or training.pop("guess")
        training.pop("stage", None)
    main_fluxes = args.flux_precision or ("fp64" if args.precision == "fp64" else "mixed")
    main_run = run_method(pcf, args, args.precision, args.preconditioner, args.initial_guess, net=net,
                          keep=True, fluxes=main_fluxes)
    st = main_run["state"]
    n_cells = args.nx * args.nx * args.nz
    prof = main_run["prof"]
    roofline = None
    if "apply_total_dot" in prof:
        launches, total_ms = prof["apply_total_dot"]
        # 25 coefficients + s + ls + the dot operand, FP64, per interior cell
        # (the operator apply is FP64 in every method)
        bytes_per_launch = 28 * 8 * n_cells
        avg_s = total_ms / launches * 1e-3
        achieved = bytes_per_launch / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": "k_stencil_pair<MODE_TOTAL, ndot> (25-point operator apply + fused "
                                              "dot products)",
                    "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": APPLY_TRAFFIC_BYTES.get((args.nx, args.nx, args.nz, "fp64")),
                    "algorithmic_bytes": bytes_per_launch,
                    "avg_launch_ms": round(total_ms / launches, 4), "launches": launches}
    # the other methods of the same workload, against the FP64 run's fields
    variants = []
    if plain and not args.no_variants:
        vnet = None
        vtrain = None
        # (preconditioner, storage of its finest-level coefficients, multigrid
        # coarsest-level steps, initial guess); the FP64 BiCGSTAB and its 1e-8
        # criterion are the same in all
        vstage = None
        # (..., whether the learned stage runs, the RK stages' flux arithmetic)
        for pc, storage, coarse, guess, staged, fluxes in (("line_relaxation", "fp32", 20, "zero", False, "fp64"),
                                                           ("line_relaxation", "bf16", 20, "zero", False, "fp64"),
                                                           ("multigrid", "fp32", 20, "zero", False, "fp64"),
                                                           ("multigrid", "bf16", 2, "zero", False, "fp64"),
                                                           ("multigrid", "bf16", 2, "previous", False, "fp64"),
                                                           ("multigrid", "bf16", 2, "previous", False, "mixed"),
                                                           ("multigrid", "bf16", 2, "extrapolated", False, "mixed"),
                                                           ("multigrid", "bf16", 2, "learned", False, "mixed"),
                                                           ("multigrid", "bf16", 2, "previous", True, "mixed")):
            if (guess == "learned" or staged) and vtrain is None:
                vtrain = train_guess(pcf, args, preconditioner=pc)
                vnet = vtrain.pop("net", None) or vtrain.pop("guess")
                vstage = vtrain.pop("stage", None)
            if staged and (vstage is None or not stage_usable(vtrain)):
                # a stage whose held-out output is not finite or blows the
                # residual up would only send every solve to the fallback
                variants.append({"method": "learned stage skipped: its held-out residual ratios "
                                           f"{vtrain.get('held_out_stage_residual_vs_multigrid')}"})
                continue
            r = run_method(pcf, args, "mixed", pc, guess, net=vnet if guess == "learned" else None, keep=True,
                           storage=storage, mg_coarse=coarse, stage_net=vstage if staged else None, fluxes=fluxes,
                           profile_apply=False)
            desc = pc.replace("_", " ") + " preconditioner"
            if pc == "multigrid":
                desc += f" ({args.mg_smoothing_steps} smoothing steps, {coarse} coarsest-level steps)"
            if staged:
                desc += (f" followed by the learned flexible-BiCGSTAB stage in the first {args.stage_iterations} "
                         "iteration(s)")
            flux_desc = "; RK-stage fluxes in FP32" if fluxes == "mixed" else ""
            entry = {"method": f"FP64 BiCGSTAB, {desc} in FP32 with {storage} coefficients, {guess} initial guess"
                               + flux_desc,
                     "value": round(r["value"], 4), "ms_per_step": round(r["ms_per_step"], 2),
                     "speedup_vs_plain": round(r["value"] / main_run["value"], 3),
                     "bicgstab_iterations_per_solve": round(r["iterations_per_solve"], 1),
                     "error_vs_fp64": field_difference(r["state"], st),
                     "kernel_ms_one_step": {k: [v[0], round(v[1], 3)] for k, v in sorted(r["breakdown"].items())},
                     "stage_kernel_shapes": r["stage_tuning"]}
            entry["setup_s"], entry["initial_cleaning_s"] = r["setup_s"], r["initial_cleaning_s"]
            if guess in ("previous", "extrapolated", "learned"):
                entry["guess_fallbacks"] = r["state"].guess_fallbacks()
            if staged:
                entry["stage_fallbacks"] = r["state"].stage_fallbacks()
            if guess == "learned" or staged:
                entry["learned_training"] = vtrain
                net = r["breakdown"].get("net")
                if net:
                    # 2 x 27 taps x (3 x 32 + 4 x 32 x 32 + 32 x 1) multiply-adds per voxel
                    flop = 2 * 27 * (3 * 32 + 4 * 32 * 32 + 32) * n_cells
                    t = net[1] / net[0] * 1e-3
                    entry["network_roofline"] = {
                        "bound": "mfma", "kernel": "network forward (k_brick: every layer an implicit GEMM on "
                                                   "bf16 MFMA over an LDS-staged input brick)",
                        "achieved": round(flop / t / 1e12, 1), "peak": 2500.0, "unit": "TFLOP/s",
                        "frac": round(flop / t / 1e12 / 2500.0, 4), "flop_per_forward": flop,
                        "avg_forward_ms": round(net[1] / net[0], 3)}
            r["state"].close()
            variants.append(entry)
    # the strict criterion |r| / |b| <= 1e-8 (tolerance_is_relative = true)
    # instead of the reference default 1e-8 / tolref
    strict = []
    if plain and not args.no_variants:
        for prec, pc, storage, coarse in (("fp64", "line_relaxation", "fp32", 20), ("mixed", "multigrid", "bf16", 2)):
            method = ("plain FP64" if prec == "fp64" else
                      f"FP64 BiCGSTAB, multigrid preconditioner in FP32 with {storage} coefficients "
                      f"({coarse} coarsest-level steps)")
            try:
                r = run_method(pcf, args, prec, pc, "zero", storage=storage, mg_coarse=coarse, relative=True)
            except pcf.PinCFlowError as e:  # e.g. no convergence within poisson_iterations
                strict.append({"method": method, "error": str(e)})
                continue
            strict.append({"method": method, "value": round(r["value"], 4), "ms_per_step": round(r["ms_per_step"], 2),
                           "bicgstab_iterations_per_solve": round(r["iterations_per_solve"], 1)})
    error_vs_fp64 = None
    fp64_run = None
    if not plain and not args.no_fp64_check:
        error_vs_fp64, fp64_run = fp64_difference(pcf, args, st)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(pcf, main_run["nl"], main_run["iters"], args)
    out = {
        "metric": f"time-steps/s at {args.nx}x{args.nx}x{args.nz}, Poisson to 1e-8 rel residual: learned vs plain",
        "value": round(main_run["value"], 4),
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if plain else ("f64 (preconditioner f32, RK-stage fluxes f32)" if main_fluxes == "mixed" else "f64 (preconditioner f32)"),
        "data": "synthetic",
        "config": {"workload": (("compressible " if args.model == "Compressible" else "pseudo-incompressible ") +
                                ("flow over 16 seeded Gaussian hills " if args.workload == "gaussian_hills"
                                 else "mountain wave ") +
                                f"{args.nx}x{args.nx}x{args.nz}, " +
                                (f"FP64 BiCGSTAB with the {args.preconditioner.replace('_', ' ')} preconditioner "
                                 f"evaluated in FP32, initial guess {args.initial_guess}, RK-stage fluxes in "
                                 f"{'FP32' if main_fluxes == 'mixed' else 'FP64'}" if not plain
                                 else "plain preconditioned BiCGSTAB (FP64)")),
                   "cells": n_cells, "poisson_tolerance": 1e-8,
                   "preconditioner_iterations": args.pc_iterations, "dtau": args.dtau,
                   **({} if plain else {"preconditioner": args.preconditioner, "storage": args.storage,
                                        "mg_smoothing_steps": args.mg_smoothing_steps,
                                        "mg_coarse_steps": args.mg_coarse_steps}),
                   "bicgstab_iterations_per_solve": round(main_run["iterations_per_solve"], 1),
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "setup_s": main_run["setup_s"],
        "initial_cleaning_s": main_run["initial_cleaning_s"],
        "variants": variants,
        "strict_criterion": strict,
        "error_vs_fp64": error_vs_fp64,
        "fp64_reference_run": fp64_run,
        "learned_training": training,
        "kernel_ms_one_step": {k: [v[0], round(v[1], 3)] for k, v in sorted(main_run["breakdown"].items())},
        "stage_kernel_shapes": main_run["stage_tuning"],
    }
    print(json.dumps(out), flush=True)


def ensemble_bench(pcf, args):
    """Member time steps per second of the seeded bubble ensemble
    (examples.bubble_member, seeds 0 .. members-1, their own grid sizes) run
    by ensemble() with all members sharing the GPU, against the same members
    one at a time; the members' records are kept in memory, not written."""
    def sim(seed):
        nl = pcf.examples.bubble_member(seed=seed, iterations=args.steps, precision=args.precision,
                                        x_size=args.member_x, z_size=args.member_z)
        if args.precision == "mixed":
            nl.poisson.preconditioner_type = args.preconditioner
            nl.poisson.preconditioner_storage = args.storage
            nl.poisson.mg_coarse_steps = args.mg_coarse_steps
            nl.discretization.flux_precision = args.flux_precision or "mixed"
        return nl
    params = {"seed": tuple(range(args.members)),
              "output_file": tuple(f"member{m}.npz" for m in range(args.members))}
    cells = [sim(m).domain for m in range(args.members)]
    cells = sum(d.x_size * d.y_size * d.z_size for d in cells)
    runs = {}
    modes = {"alone": ("one at a time", 1, False), "streams": ("sharing the GPU", args.members, False),
             "batched": ("sharing the GPU, batched solves", args.members, True)}
    for mode in args.ensemble_modes.split(","):
        label, conc, batched = modes[mode]
        batch = pcf.SolveBatch() if batched else None
        t0 = time.perf_counter()
        def progress(member, itime, info, label=label, conc=conc):
            if (member == 1 and conc > 1) or (conc == 1 and itime % 10 == 0):
                print(f"# [ensemble] {label}: member {member} at step {itime}", file=sys.stderr, flush=True)
        res = pcf.ensemble(sim, params, {}, max_concurrent=conc, write=False, batched=batched, batch=batch,
                           on_step=progress)
        elapsed = time.perf_counter() - t0
        its = [sum(s[2] + s[3] for s in r[2]) / (2 * len(r[2])) for r in res]
        for r in res:
            r[0].close()
        runs[label] = {"member_steps_per_s": round(args.members * args.steps / elapsed, 3),
                       "wall_s": round(elapsed, 2), "bicgstab_iterations_per_solve": round(sum(its) / len(its), 1)}
        if batch is not None:
            runs[label]["batch"] = batch.stats()
            batch.close()
        print(f"# [ensemble] {label}: {elapsed:.1f}s", file=sys.stderr, flush=True)
    base = runs.get("one at a time", runs.get("sharing the GPU"))
    sizes = (f", every member {args.member_x}x{args.member_x}x{args.member_z}"
             if args.member_x and args.member_z else "")
    print(json.dumps({"workload": f"seeded bubble ensemble, {args.members} members, {args.steps} steps each "
                                  f"(setup and initial cleaning included){sizes}",
                      "cells": cells, "precision": args.precision,
                      **({"preconditioner": args.preconditioner,
                          "flux_precision": args.flux_precision or "mixed"} if args.precision == "mixed" else {}),
                      "runs": runs,
                      "speedup": {k: round(v["member_steps_per_s"] / base["member_steps_per_s"], 3)
                                  for k, v in runs.items()}}), flush=True)


def train_guess(pcf, args, preconditioner=None):
    """Train the initial-guess network on the solver's own solves of
    held-out configurations (the same physics with seeded random hills, not
    the benchmark's single hill), on the device; returns the inference
    network and a summary (not part of the timed region)."""
    if args.train_loss == "residual":
        return train_guess_residual(pcf, args, preconditioner)
    t0 = time.perf_counter()
    samples = []
    for r in range(args.train_runs):
        nl = pcf.examples.gaussian_hills_3d(x_size=args.nx, y_size=args.nx, z_size=args.nz,
                                            iterations=args.train_steps, seed=100 + r, n_hills=8, precision="mixed")
        nl.poisson.preconditioner_type = preconditioner or args.preconditioner
        nl.atmosphere.model = args.model
        st = pcf.State(nl)
        st.initial_cleaning()
        samples += pcf.learned.collect_samples(st, args.train_steps)
        st.close()
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    net = pcf.learned.GuessNetwork(args.nx, args.nx, args.nz, training=True, seed=0)
    losses = pcf.learned.train(net, samples, epochs=args.train_epochs,
                               log=lambda m: print("# " + m, file=sys.stderr, flush=True))
    params = net.params()
    net.close()
    t_train = time.perf_counter() - t0
    infer = pcf.learned.GuessNetwork(args.nx, args.nx, args.nz, training=False, params=params)
    # squared error of the guess relative to the warm start's on the
    # benchmark's own terrain, which the training runs do not contain
    nl = pcf.examples.mountain_wave_3d(x_size=args.nx, y_size=args.nx, z_size=args.nz, iterations=4,
                                       precision="mixed")
    nl.poisson.preconditioner_type = preconditioner or args.preconditioner
    nl.atmosphere.model = args.model
    st = pcf.State(nl)
    st.initial_cleaning()
    held_out = pcf.learned.evaluate(infer, pcf.learned.collect_samples(st, 4))
    st.close()
    return {"net": infer, "samples": len(samples), "epochs": args.train_epochs, "loss_first": losses[0],
            "loss_last": losses[-1], "held_out_error_vs_warm_start": held_out, "data_s": round(t_data, 1),
            "train_s": round(t_train, 1)}


def stage_usable(summary) -> bool:
    r = summary.get("held_out_stage_residual_vs_multigrid") or []
    return bool(r) and all(math.isfinite(x) and x < 10.0 for x in r)


def train_guess_residual(pcf, args, preconditioner=None):
    """Online training on the device (learned.train_online) over the solves
    of seeded random-hill runs: the initial-guess network and the network of
    the flexible-BiCGSTAB preconditioning stage, then, on the benchmark's own
    terrain (which the training runs do not contain), the residual of each
    relative to what it corrects (the warm start, M^-1 of the first
    residual).  Returns both inference networks and a summary."""
    t0 = time.perf_counter()
    g = pcf.learned.GuessNetwork(args.nx, args.nx, args.nz, training=True, seed=0)
    s = pcf.learned.GuessNetwork(args.nx, args.nx, args.nz, training=True, seed=1)
    hist = []
    log = (lambda m: print("# " + m, file=sys.stderr, flush=True))
    for r in range(args.train_runs):
        nl = pcf.examples.gaussian_hills_3d(x_size=args.nx, y_size=args.nx, z_size=args.nz,
                                            iterations=args.train_steps, seed=100 + r, n_hills=8, precision="mixed")
        nl.poisson.preconditioner_type = preconditioner or args.preconditioner
        nl.poisson.preconditioner_storage = args.storage
        nl.poisson.mg_coarse_steps = 2
        nl.atmosphere.model = args.model
        st = pcf.State(nl)
        st.initial_cleaning()
        hist += pcf.learned.train_online([st], args.train_steps, guess=g, stage=s, inner=args.train_inner,
                                         lr=args.train_lr, objective=args.train_objective, log=log)
        st.close()
    gp, sp = g.params(), s.params()
    g.close()
    s.close()
    t_train = time.perf_counter() - t0
    guess = pcf.learned.GuessNetwork(args.nx, args.nx, args.nz, training=False, params=gp)
    stage = pcf.learned.GuessNetwork(args.nx, args.nx, args.nz, training=False, params=sp)
    # held out: the benchmark's terrain, evaluated without training (inner = 0)
    nl = pcf.examples.mountain_wave_3d(x_size=args.nx, y_size=args.nx, z_size=args.nz, iterations=5,
                                       precision="mixed")
    nl.poisson.preconditioner_type = preconditioner or args.preconditioner
    nl.poisson.preconditioner_storage = args.storage
    nl.poisson.mg_coarse_steps = 2
    nl.atmosphere.model = args.model
    st = pcf.State(nl)
    st.initial_cleaning()
    held = pcf.learned.train_online([st], 5, guess=guess, stage=stage, inner=0)
    st.close()

    def ratios(recs, key):
        return [round(r[key][1] / r[key][0], 4) for r in recs]

    return {"guess": guess, "stage": stage, "objective": args.train_objective, "solves": len(hist),
            "adam_steps_per_solve": args.train_inner, "learning_rate": args.train_lr, "train_s": round(t_train, 1),
            "training_guess_residual_vs_warm_start_last5": ratios(hist[-5:], "guess"),
            "training_stage_residual_vs_multigrid_last5": ratios(hist[-5:], "stage"),
            "held_out_guess_residual_vs_warm_start": ratios(held, "guess"),
            "held_out_stage_residual_vs_multigrid": ratios(held, "stage")}


def fp64_difference(pcf, args, st_mixed):
    """Max-norm differences, relative to the field's max norm, between the
    mixed-precision fields after warmup + steps and an FP64 run of the same
    steps (the FP64 library path agrees with the FP64 oracle to 1e-9); the
    Exner pressure is compared up to a constant."""
    nl = workload(pcf, args, args.steps + args.warmup + 1, "fp64")
    ref = pcf.State(nl)
    ref.initial_cleaning()
    t0 = time.perf_counter()
    its = []
    for s in range(args.steps + args.warmup + 1):
        info = ref.step()
        its += list(info.niter)
        print(f"# [fp64 check] step {s + 1}: iterations {list(info.niter)}", file=sys.stderr, flush=True)
    fp64_rate = (args.steps + args.warmup + 1) / (time.perf_counter() - t0)
    out = {}
    for f in ("RHO", "RHOP", "U", "V", "W", "PIP"):
        a, b = st_mixed.get(f), ref.get(f)
        d = a - b
        if f == "PIP":  # up to a constant: the null space of the pressure operator
            d = d - d[3:-3, 3:-3, 3:-3].mean()
        scale = float(abs(b).max()) or 1.0
        out[f] = float(abs(d).max()) / scale
    ref.close()
    return out, {"value": round(fp64_rate, 4), "bicgstab_iterations_per_solve": round(sum(its) / len(its), 1),
                 "steps": args.steps + args.warmup + 1}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_components(pcf, oracle, nl, sx, threads):
    """Host seconds of the components of one step of the FP64 oracle on sx x
    sx columns of the workload: explicit LHS stage, RHS stage, RHS + operator
    assembly, and one BiCGSTAB iteration (10 timed, the solve stopped there)."""
    sub = pcf.examples.mountain_wave_3d(x_size=sx, y_size=sx, z_size=nl.domain.z_size, iterations=1,
                                        precision="fp64")
    sub.atmosphere.model = nl.atmosphere.model
    cfg = pcf.to_config(sub)
    cfg.poisson_iterations = 10
    o = oracle.Oracle(cfg, pcf._abi, threads=threads)
    o.compute_sponges(0.01, 0.1)
    o.backup_predictands()
    t0 = time.perf_counter()
    o.explicit_integration_lhs(0.01, 0.1)
    t_lhs = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.explicit_integration_rhs(0.01, 0.1)
    t_rhs = time.perf_counter() - t0
    t0 = time.perf_counter()
    tolref = o.compute_lhs()
    o.compute_operator(0.01, 1.0)
    t_op = time.perf_counter() - t0
    t0 = time.perf_counter()
    rc, n_done = o.apply_bicgstab(tolref)
    t_it = (time.perf_counter() - t0) / max(n_done, 1)
    return {"lhs": t_lhs, "rhs": t_rhs, "assembly": t_op, "iteration": t_it, "timed_iterations": n_done}


def cpu_baseline(pcf, nl, iters, args):
    """The FP64 oracle on the host's cores, timed on a bounded sample of the
    same workload.  The components of one step (explicit LHS stage, RHS
    stage, RHS + operator assembly, one BiCGSTAB iteration) are timed on a
    horizontal subset of the columns and scaled to the full grid by the cell
    count and to a step by the GPU run's iteration count.  The model is
    checked in the same run against one whole oracle step (integrate's time
    loop body, both solves to the reference criterion) timed on a smaller
    sample, where the model uses that step's own iteration counts.  Threads:
    OMP_NUM_THREADS if set (the host's share of cores for one GPU), else all
    cores."""
    from oracle import oracle

    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = args.cpu_threads or (int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1))
    sx = min(args.cpu_sample_nx, nl.domain.x_size)
    nz = nl.domain.z_size
    full_cells = nl.domain.x_size * nl.domain.y_size * nz
    scale = full_cells / (sx * sx * nz)
    c = oracle_components(pcf, oracle, nl, sx, threads)
    n_it = sum(sum(i) for i in iters) / len(iters)
    t_step = scale * (2 * c["lhs"] + c["rhs"] + 2 * c["assembly"] + n_it * c["iteration"])
    # the model against a whole timed oracle step on sc x sc columns
    scx = min(64, sx)
    cc = oracle_components(pcf, oracle, nl, scx, threads)
    whole = pcf.examples.mountain_wave_3d(x_size=scx, y_size=scx, z_size=nz, iterations=1, precision="fp64")
    whole.atmosphere.model = nl.atmosphere.model
    o = oracle.Oracle(pcf.to_config(whole), pcf._abi, threads=threads)
    o.initial_cleaning()
    t0 = time.perf_counter()
    info = o.step()
    t_whole = time.perf_counter() - t0
    n_whole = info.niter[0] + info.niter[1]
    t_model = 2 * cc["lhs"] + cc["rhs"] + 2 * cc["assembly"] + n_whole * cc["iteration"]
    return {"value": round(1.0 / t_step, 6), "unit": "time-steps/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": (f"FP64 oracle on {sx}x{sx}x{nz} columns of the same workload: explicit LHS stage "
                       f"{c['lhs']:.2f}s, RHS stage {c['rhs']:.2f}s, RHS+operator assembly {c['assembly']:.2f}s, "
                       f"{c['iteration'] * 1e3:.1f} ms per BiCGSTAB iteration ({c['timed_iterations']} timed); "
                       f"step = 2 LHS + RHS + 2 assemblies + {n_it:.0f} iterations (the GPU run's count), scaled "
                       f"x{scale:.0f} to the full grid"),
            "s_per_step": round(t_step, 1),
            "model_check": {"sample": f"{scx}x{scx}x{nz}", "timed_step_s": round(t_whole, 3),
                            "model_step_s": round(t_model, 3), "iterations": [info.niter[0], info.niter[1]]}}


if __name__ == "__main__":
    main()
