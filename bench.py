#!/usr/bin/env python3
"""Benchmark: load-steps/s of the quasi-static compression solve on one GPU.

Workload (the load-steps/s metric's configuration): simple cubic lattice
32x32x16 cells (18,513 nodes, 53,328 beams), cell 5 mm, radius graded
0.4 -> 0.6 mm along the load axis, E 2.0 GPa, nu 0.35, yield 50 MPa, linear
hardening 5% E, platen to 30% strain in 60 load + 60 unload increments,
L-BFGS m = 10, absolute gradient tolerance 1e-6, self-contact on.

A "step" is one scheduled load increment (one equilibrium solve, bisection
sub-steps included).  The solver state starts from the undeformed lattice on
the device; --warmup increments (the tare first) run untimed, then exactly
--steps increments are timed between device synchronisations.

The JSON line also carries:
  roofline      the dominant kernel's algorithmic bytes per launch divided by
                its mean launch duration, measured with HIP events on the solver
                stream during the timed region (peak: MI355X HBM 8 TB/s)
  cpu_baseline  the FP64 CPU oracle (single-threaded restatement of the
                reference solver) on the host, on a bounded sample of the same
                workload (the tare and the first load increments)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "falling-dice_amd"))

HBM_PEAK_GBS = 8000.0


def workload(args):
    import mmc
    cfg = mmc.cubic_config(args.nx, args.ny, args.nz, load=args.load_steps, unload=args.unload_steps)
    name = (f"cubic {args.nx}x{args.ny}x{args.nz} cells, cell 5 mm, r 0.4->0.6 mm, "
            f"E 2.0 GPa, yield 50 MPa, H 5%E, 30% strain, {args.load_steps}+{args.unload_steps} increments, "
            f"L-BFGS m=10, grad tol 1e-6")
    return cfg, name


def beam_kernel_bytes(N, B, precision):
    """Compulsory HBM bytes of one k_beam launch: per beam the endpoints (8),
    L0/radius/area/I (32), director (24), plastic state (56) and the 9-double
    gradient record written (72); per node the pose read once (x 24 B, R 32 B
    FP64 or 16 B FP32)."""
    rbytes = 32 if precision == 0 else 16
    return B * (8 + 32 + 24 + 56 + 72) + N * (24 + rbytes)


def cpu_baseline(cfg, budget_s):
    sys.path.insert(0, str(ROOT / "oracle"))
    import py_oracle
    o = py_oracle.Solver(cfg, "orc")
    z = np.zeros(6 * o.n_nodes)
    t0 = time.perf_counter()
    steps = 0
    # tare (untimed in both legs: it converges in one iteration), then load increments
    o.minimize(o.top0, z)
    o.commit()
    t0 = time.perf_counter()
    while steps < cfg.load_steps:
        steps += 1
        top = o.top0 - cfg.max_strain * o.height * steps / cfg.load_steps
        st, fin, _ = o.minimize(top)
        o.commit()
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": steps / dt, "unit": "load-steps/s", "cores": 1, "kind": "port",
            "sample": f"load increments 1..{steps} after the tare ({dt:.1f} s on one host core)",
            "steps": steps, "seconds": dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed increments (0 = rest of the schedule)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed increments before timing (tare first)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--nx", type=int, default=32)
    ap.add_argument("--ny", type=int, default=32)
    ap.add_argument("--nz", type=int, default=16)
    ap.add_argument("--load-steps", type=int, default=60)
    ap.add_argument("--unload-steps", type=int, default=60)
    ap.add_argument("--kernel", default="beam", help="kernel whose roofline is reported")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this solver has no multi-GPU path (pass --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import mmc
    prec = mmc.MMC_FP64 if args.precision == "fp64" else mmc.MMC_MIXED
    cfg, name = workload(args)
    sim = mmc.Simulator(cfg, prec)
    N, B = sim.n_nodes, sim.n_beams
    total = 1 + args.load_steps + args.unload_steps
    warm = max(0, min(args.warmup, total - 1))
    K = args.steps if args.steps > 0 else total - warm
    K = min(K, total - warm)
    sim.begin()
    frames = []
    for _ in range(warm):
        f, _ = sim.advance()
        frames.append(f)
        print(f"warmup increment done ({len(frames)}/{warm})", file=sys.stderr, flush=True)
    ev0 = sim.evaluations()
    t0 = time.perf_counter()
    for k in range(K):
        f, _ = sim.advance()
        frames.append(f)
        if k % 10 == 9:
            print(f"timed increment {k + 1}/{K}", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    evals = sim.evaluations() - ev0
    # the dominant kernel, timed with HIP events on the solver stream at the
    # state the timed region ended in (launches inside graph replays carry no
    # per-launch events)
    launches = 200
    mean_ms = sim.time_kernel(args.kernel, launches)
    others = {k: round(sim.time_kernel(k, 100) * 1e3, 3) for k in ("state", "contact", "gather") if k != args.kernel}
    fr = np.concatenate(frames)
    value = K / dt
    bytes_launch = beam_kernel_bytes(N, B, prec) if args.kernel == "beam" else None
    roof = None
    if bytes_launch and mean_ms > 0:
        ach = bytes_launch / (mean_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": f"k_beam",
                "bytes_per_launch": bytes_launch, "mean_launch_us": round(mean_ms * 1e3, 3),
                "launches_timed": launches, "other_kernels_us": others}
    out = {
        "metric": "load-steps/sec, mixed vs FP64 on 1 GPU, at <=1e-5 rel platen-force error",
        "value": round(value, 4), "unit": "load-steps/s", "n_gpus": 1, "steps": K, "warmup": warm,
        "ms_per_step": round(dt / K * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if prec == 0 else "f32-element/f64-reduction",
        "data": "synthetic (seeded cubic lattice)",
        "config": {"workload": name, "nodes": N, "beams": B, "precision": args.precision,
                   "increments": f"{warm} warmup + {K} timed of {total}"},
        "lbfgs_iterations": int(fr["iterations"][warm:].sum()) if len(fr) > warm else None,
        "objective_evaluations": int(evals),
        "beam_evaluations_per_s": round(evals * B / dt, 1),
        "peak_force_n": float(fr["reaction_force"].max()),
        "all_converged": bool(np.all(fr["converged"] == 1)),
        "roofline": roof,
    }
    if not args.no_cpu:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
