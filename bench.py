#!/usr/bin/env python3
"""Benchmark: load-steps/s of the quasi-static compression solve on one GPU,
mixed precision against plain FP64, at <= 1e-5 relative platen-force error.

Workload: simple cubic lattice 32x32x16 cells (18,513 nodes, 53,328 beams),
cell 5 mm, radius graded 0.4 -> 0.6 mm along the load axis, E 2.0 GPa, nu 0.35,
yield 50 MPa, linear hardening 5% E, platen to 30% strain in 60 load + 60
unload increments, L-BFGS m = 10, self-contact on: the 32x32x16 configuration
of the metric, in the parity form the error qualifier needs: absolute gradient
tolerance 1e-7 (tighter than the configuration's 1e-6, which leaves two
correct solvers ~1e-5 apart) with the per-term Armijo test and its
approximate-Wolfe completion (line_search 2), 20,000 iterations per solve.

A "step" is one scheduled load increment (one equilibrium solve, bisection
sub-steps included).  Each leg builds its solver on the device; --warmup
increments (the tare first) run untimed, then exactly --steps increments are
timed between device synchronisations.  The mixed leg (FP32 rotational log map
and moment transport, FP64 elsewhere) gives `value`; the FP64 leg runs the same
increments.  Both legs' platen-force curves are compared with each other and
with the oracle's FP64 curve of the same workload (tests/golden), frame by
frame over every frame both cover: `qualifier` states whether the metric's
error bound holds (mixed <= 1e-5 against FP64 and against the oracle) and,
separately, whether the FP64 leg is within its own 1e-6 of the oracle.

The JSON line also carries:
  roofline      the device loop's dominant kernel (k_gather_g: gradient
                assembly and curvature-pair products of a trial): algorithmic
                bytes per launch (200 B per DOF at a full history) over its
                mean launch duration, measured with HIP events on the solver
                stream at the state the timed region ended in (peak: MI355X HBM
                8 TB/s); `loop_kernel_us` lists the four loop kernels' mean
                durations measured the same way
  cpu_baseline  the FP64 CPU oracle (single-threaded restatement of the
                reference solver) on one host core, on a bounded sample of the
                same increments: its L-BFGS loop on the first timed increment,
                started from the FP64 leg's committed state after the warmup,
                capped at about --cpu-budget seconds, scaled to the window's
                L-BFGS iterations
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
# HBM bytes per k_gather_g launch from rocprofv3 FETCH_SIZE + WRITE_SIZE passes on
# the timed window, calibrated against a device copy of known size (commands
# and figures in profiles/); keyed by lattice cells and gradient tolerance
GATHER_TRAFFIC_BYTES: dict = {
    # 32x32x16, gradient tolerance 1e-7, mixed policy, at a full history:
    # FETCH_SIZE 16,126.4 KiB x 1.999 + WRITE_SIZE 3,127.6 KiB x 1.000
    # (profiles/k_gather_g_cubic32x32x16.md)
    (32, 32, 16, 1e-7, "mixed"): 36219501,
}


def workload(args):
    import mmc
    if args.workload == "large":
        # contact-heavy lattice: 64x64x32 cells, node positions jittered by a
        # seeded Gaussian of 2% of the cell, per-beam radii uniform 0.3..0.7 mm,
        # platen to 50% strain in 100 + 100 increments
        cell = 0.005
        cfg = mmc.cubic_config(args.nx, args.ny, args.nz, cell=cell, r_bot=3e-4, r_top=7e-4, strain=0.5,
                               load=100, unload=100, grad_tol=args.grad_tol, max_iterations=args.max_iterations,
                               sigma=0.02 * cell, radius_mode=1)
        # the same line-search form as the 32x32x16 workload: the reference's
        # total-energy Armijo test cannot resolve decreases at tolerance 1e-7
        cfg = cfg.replace(line_search=args.line_search)
        name = (f"cubic {args.nx}x{args.ny}x{args.nz} cells, cell 5 mm, node jitter N(0, (0.1 mm)^2), "
                f"per-beam r U(0.3, 0.7) mm, E 2.0 GPa, yield 50 MPa, H 5%E, 50% strain, 100+100 increments, "
                f"L-BFGS m=10, grad tol {args.grad_tol:g}"
                + {0: "", 1: ", per-term Armijo", 2: ", per-term Armijo + approximate Wolfe (parity form)"}[args.line_search]
                + f", max {args.max_iterations} iterations")
        return cfg, name
    cfg = mmc.cubic_config(args.nx, args.ny, args.nz, load=args.load_steps, unload=args.unload_steps,
                           grad_tol=args.grad_tol, strain=args.strain, max_iterations=args.max_iterations)
    cfg = cfg.replace(line_search=args.line_search)
    name = (f"cubic {args.nx}x{args.ny}x{args.nz} cells, cell 5 mm, r 0.4->0.6 mm, "
            f"E 2.0 GPa, yield 50 MPa, H 5%E, {args.strain * 100:g}% strain, "
            f"{args.load_steps}+{args.unload_steps} increments, L-BFGS m=10, grad tol {args.grad_tol:g}"
            + {0: "", 1: ", per-term Armijo", 2: ", per-term Armijo + approximate Wolfe (parity form)"}[args.line_search]
            + f", max {args.max_iterations} iterations")
    return cfg, name


def oracle_fixture(args):
    """The oracle's FP64 curve of this workload, when tests/golden holds it."""
    if (args.nx, args.ny, args.nz, args.strain, args.load_steps, args.unload_steps) != (32, 32, 16, 0.30, 60, 60):
        return None, None
    name = {(1e-7, 2): "cubic_32x32x16_tight", (1e-6, 0): "cubic_32x32x16"}.get((args.grad_tol, args.line_search))
    path = ROOT / "tests" / "golden" / f"{name}_oracle_frames.npy"
    return (np.load(path), name) if name and path.exists() else (None, None)


def gather_loop_bytes(N, B, m):
    """Algorithmic HBM bytes of one k_gather_g launch at a full history (m - 1
    retained pairs).  Per DOF element (6N of them): z, z_trial, g and the
    direction p read and the trial gradient written (5 doubles), the m - 1
    retained s_j, y_j read for the pair products and the new pair s, y written
    (2m doubles).  Per beam: its two end records read by the gather (2 x 6
    doubles).  At m = 10: 200 B per DOF plus 96 B per beam."""
    return 6 * N * (5 + 2 * m) * 8 + 2 * B * 6 * 8


def run_leg(cfg, prec, warmup, steps, log, keep_state=False):
    """One precision leg: warmup increments untimed, then `steps` timed.
    keep_state: also return the committed state after the warmup (x, q,
    plastic), the starting point of the CPU baseline's sample."""
    import mmc
    sim = mmc.Simulator(cfg, prec)
    total = 1 + cfg.load_steps + cfg.unload_steps
    warm = max(0, min(warmup, total - 1))
    K = steps if steps > 0 else total - warm
    K = min(K, total - warm)
    label = "mixed" if prec else "fp64"
    sim.begin()
    frames = []
    for _ in range(warm):
        f, _ = sim.advance()
        frames.append(f)
    log(f"{label}: {warm} warmup increments done")
    state = sim.read_state() if keep_state else None
    ev0 = sim.evaluations()
    t0 = time.perf_counter()
    marks = []
    for k in range(K):
        f, _ = sim.advance()
        frames.append(f)
        marks.append(time.perf_counter())
        log(f"{label}: timed increment {k + 1}/{K} ({int(f['iterations'].sum())} iterations)")
    dt = time.perf_counter() - t0
    evals = sim.evaluations() - ev0
    fr = np.concatenate(frames)
    # wall time of each timed increment (the host waits for every increment's
    # status, so consecutive marks bracket one increment's device work)
    sim.increment_s = np.diff(np.concatenate([[t0], marks]))
    sim.warm_state = state
    return sim, fr, warm, K, dt, evals


def first_instability(*legs, jump=4.0):
    """First frame at which any leg's L-BFGS iteration count exceeds `jump`
    times the largest count of its earlier frames (the tare, frame 0, aside).
    Such a jump marks the loss of stability of the equilibrium path (the
    Hessian turns near-singular where the lattice starts to buckle); from
    there on the solution branch is selected by round-off, so two solvers
    that round differently need not follow the same branch."""
    first = None
    for f in legs:
        it = f["iterations"]
        for k in range(2, len(it)):
            if it[k] > jump * max(1, int(it[1:k].max())):
                first = k if first is None else min(first, k)
                break
    return first


def curve_error(fa, fb):
    n = min(len(fa), len(fb))
    if n == 0:
        return None
    peak = max(float(np.max(fb["reaction_force"][:n])), 1e-12)
    d = np.abs(fa["reaction_force"][:n] - fb["reaction_force"][:n]) / peak
    kp = int(np.argmax(fb["reaction_force"][:n]))
    # frames before the first instability, where the equilibrium path is unique
    ku = first_instability(fa[:n], fb[:n])
    ku = n if ku is None else ku
    return {"frames": int(n), "max_rel_force": float(d.max()), "worst_frame": int(d.argmax()),
            "unique_path_frames": int(ku), "max_rel_force_unique_path": float(d[:ku].max()),
            "unique_path_strain": float(fb["engineering_strain"][ku - 1]),
            "peak_frame": kp, "max_rel_force_to_peak": float(d[:kp + 1].max()),
            "loading_work_rel": float(abs(fa["loading_work"][n - 1] - fb["loading_work"][n - 1])
                                      / max(abs(fb["loading_work"][n - 1]), 1e-12)),
            "permanent_set_abs": float(abs(fa["permanent_set"][n - 1] - fb["permanent_set"][n - 1]))}


def run_batch(args, log):
    """Batched workloads: lattices x load-steps per second over whole schedules."""
    import mmc
    kind = args.workload
    count = args.lattices or (256 if kind == "uneven" else 1024)
    cfgs = mmc.batch_workload(kind, count, args.load_steps, args.unload_steps)
    prec = mmc.MMC_FP64 if args.precision == "fp64" else mmc.MMC_MIXED
    if args.shards:
        os.environ["MMC_BATCH_SHARDS"] = str(args.shards)
    t0 = time.perf_counter()
    sim = mmc.BatchSimulator(cfgs, prec, device_lattices=not args.host_lattices)
    setup = time.perf_counter() - t0
    source = "host" if args.host_lattices else "device"
    log(f"batch of {count}: {sum(sim.n_nodes)} nodes, {sum(sim.n_beams)} beams (lattices built on the {source}, "
        f"setup {setup:.2f} s)")
    t0 = time.perf_counter()
    frames, status, err = sim.run(verbose=True)
    dt = time.perf_counter() - t0
    steps = 1 + args.load_steps + args.unload_steps
    ok = [b for b in range(count) if status[b] == 0]
    out = {
        "metric": "lattices x load-steps/sec (batched)", "value": round(len(ok) * steps / dt, 3),
        "unit": "lattice-load-steps/s", "n_gpus": 1, "higher_is_better": True,
        "designs_per_s": round(len(ok) / dt, 4), "seconds": round(dt, 2), "setup_s": round(setup, 3),
        "lattices_built_on": source, "shards": int(os.environ.get("MMC_BATCH_SHARDS", "1")),
        "dtype": "f64" if prec == 0 else "f32-element/f64-reduction", "data": "synthetic (seeded lattices)",
        "config": {"workload": kind, "lattices": count, "nodes": int(sum(sim.n_nodes)),
                   "beams": int(sum(sim.n_beams)), "increments": f"{args.load_steps}+{args.unload_steps} + tare"},
        "completed": len(ok), "failed": count - len(ok), "first_failure": err or None,
        "evaluations": sim.evaluations(),
    }
    if args.save_frames:
        np.save(f"{args.save_frames}_{kind}_{args.precision}.npy", np.concatenate(frames), allow_pickle=False)
    sim.close()
    rng = np.random.default_rng(7)
    if args.single_sample:
        # the plain FP64 single-lattice GPU solver, one lattice after another
        pick = sorted(rng.choice(ok, size=min(args.single_sample, len(ok)), replace=False).tolist())
        t0 = time.perf_counter()
        failed = []
        for b in pick:
            try:
                mmc.Simulator(cfgs[b], mmc.MMC_FP64).run()
            except mmc.MMCError:
                failed.append(b)  # an equilibrium failure of the FP64 solver; its time still counts
        ts = time.perf_counter() - t0
        single = (len(pick) - len(failed)) * steps / ts
        out["fp64_single_lattice"] = {"lattices": pick, "failed": failed, "seconds": round(ts, 2),
                                      "value": round(single, 3), "unit": "lattice-load-steps/s"}
        out["speedup_vs_fp64_single_lattice"] = round(out["value"] / single, 2)
        log(f"single-lattice FP64 sample done ({ts:.1f} s)")
    if args.oracle_sample:
        # per-lattice dissipated-work error against the FP64 oracle on a seeded
        # sample, and the oracle's own rate on it (one host core)
        sys.path.insert(0, str(ROOT / "oracle"))
        import py_oracle
        pick = sorted(rng.choice(ok, size=min(args.oracle_sample, len(ok)), replace=False).tolist())
        errs, failed = [], {}
        t0 = time.perf_counter()
        for b in pick:
            try:
                fo = py_oracle.Solver(cfgs[b], "orc").run()
            except RuntimeError as e:
                # the oracle's own schedule can end in an equilibrium failure
                # where the device run (other summation order) went through:
                # past buckling the two take different branches
                failed[int(b)] = str(e)
                continue
            fg = frames[b]
            if len(fo) == len(fg):
                wo, wg = fo["dissipated_work"][-1], fg["dissipated_work"][-1]
                errs.append(abs(wg - wo) / max(abs(wo), 1e-12))
        tc = time.perf_counter() - t0
        out["oracle_sample"] = {"lattices": pick, "max_rel_dissipated_work": max(errs) if errs else None,
                                "compared": len(errs), "oracle_failed": failed or None}
        out["cpu_baseline"] = {"value": round(len(pick) * steps / tc, 4), "unit": "lattice-load-steps/s",
                               "cores": 1, "kind": "port",
                               "sample": f"{len(pick)} lattices of the set, full schedules ({tc:.1f} s, one core)"}
        log(f"oracle sample done ({tc:.1f} s)")
    print(json.dumps(out))


def reference_speed_ratio():
    """The oracle's wall time over the compiled reference's on identical work,
    recorded in the build container (tests/golden/make_speed_ratio.py)."""
    path = ROOT / "tests" / "golden" / "oracle_speed_ratio.json"
    if not path.exists():
        return None
    cases = json.loads(path.read_text())["cases"]
    return {c["case"]: c["oracle_over_reference"] for c in cases}


def cpu_baseline_sample(cfg, state, top, budget_s, gpu_iterations, steps, gpu_seconds):
    """The oracle (one host core) on the first timed increment, from the FP64
    leg's committed state after the warmup: its L-BFGS loop capped so that it
    runs about budget_s, seconds per iteration from that, scaled to the GPU
    window's iteration count."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import py_oracle
    probe = py_oracle.Solver(cfg, "orc")
    x, q, p = state
    probe.set_pose(x, q)
    probe.set_plastic(p)
    z = np.zeros(6 * probe.n_nodes)
    t0 = time.perf_counter()
    probe.objective(z, top)
    t_eval = time.perf_counter() - t0
    probe.close()
    cap = max(2, int(budget_s / max(1.3 * t_eval, 1e-6)))
    o = py_oracle.Solver(cfg.replace(max_iterations=cap), "orc")
    o.set_pose(x, q)
    o.set_plastic(p)
    t0 = time.perf_counter()
    st, _, _ = o.minimize(top)
    dt = time.perf_counter() - t0
    per_it = dt / max(1, int(st.iterations))
    proj = per_it * gpu_iterations
    return {"value": round(steps / proj, 6), "unit": "load-steps/s", "cores": 1, "kind": "port",
            "sample": f"the first timed increment (platen {top:.6f} m) from the FP64 leg's committed state after the "
                      f"warmup, the oracle's first {st.iterations} L-BFGS iterations ({st.evaluations} objective "
                      f"evaluations, {dt:.1f} s on one host core, {per_it * 1e3:.2f} ms per iteration), scaled to "
                      f"the {gpu_iterations} iterations of the {steps} timed increments",
            "seconds_per_iteration": round(per_it, 6), "projected_seconds": round(proj, 1),
            "oracle_over_reference_time": reference_speed_ratio(),
            "gpu_seconds_per_iteration": round(gpu_seconds / max(1, gpu_iterations), 8),
            "gpu_over_cpu_per_iteration": round(per_it / max(gpu_seconds / max(1, gpu_iterations), 1e-12), 1)}


def cpu_baseline_window(cfg, budget_s, iterations_window, steps_window):
    """CPU rate for a workload whose increments take the oracle minutes each:
    the oracle's L-BFGS loop on load increment 1, capped so that it runs for
    about budget_s, gives seconds per iteration (line-search trials included);
    the timed GPU window's iteration count turns that into the oracle's time
    for the same increments."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import py_oracle
    probe = py_oracle.Solver(cfg, "orc")
    z = np.zeros(6 * probe.n_nodes)
    t0 = time.perf_counter()
    probe.objective(z, probe.top0)
    t_eval = time.perf_counter() - t0
    probe.close()
    cap = max(2, int(budget_s / max(1.3 * t_eval, 1e-6)))
    o = py_oracle.Solver(cfg.replace(max_iterations=cap), "orc")
    top = o.top0 - cfg.max_strain * o.height / cfg.load_steps
    t0 = time.perf_counter()
    st, _, _ = o.minimize(top, np.zeros(6 * o.n_nodes))
    dt = time.perf_counter() - t0
    per_it = dt / max(1, int(st.iterations))
    proj = per_it * iterations_window
    return {"value": round(steps_window / proj, 6), "unit": "load-steps/s", "cores": 1, "kind": "port",
            "sample": f"platen of load increment 1 from the undeformed lattice, first {st.iterations} L-BFGS "
                      f"iterations ({st.evaluations} objective evaluations, {dt:.1f} s on one host core, "
                      f"{per_it * 1e3:.1f} ms per iteration), scaled to the {iterations_window} iterations of "
                      f"the timed increments",
            "seconds_per_iteration": round(per_it, 5), "projected_seconds": round(proj, 1)}


def qualifier(mixed_fr, fp64_fr, fo, fixture):
    """The metric's error bound over every frame the legs and the oracle's
    curve share: the mixed leg (the value) within 1e-5 of the peak force
    against FP64 and against the oracle (`met`); the FP64 path's own parity
    bound, 1e-6 against the oracle, reported beside it (`fp64_within_1e-6`)."""
    q = {"mixed_vs_fp64": curve_error(mixed_fr, fp64_fr)}
    if fo is None:
        q.update(met=None, reason="no oracle curve of this workload in tests/golden")
        return q
    q["oracle_curve"] = f"tests/golden/{fixture}_oracle_frames.npy ({len(fo)} frames)"
    q["mixed_vs_oracle"] = curve_error(mixed_fr, fo)
    q["fp64_vs_oracle"] = curve_error(fp64_fr, fo)
    n = min(len(mixed_fr), len(fp64_fr), len(fo))
    q["frames_compared"] = int(n)
    q["met"] = bool(q["mixed_vs_fp64"]["max_rel_force"] <= 1e-5 and q["mixed_vs_oracle"]["max_rel_force"] <= 1e-5
                    and n == len(mixed_fr) == len(fp64_fr))
    q["fp64_within_1e-6"] = bool(q["fp64_vs_oracle"]["max_rel_force"] <= 1e-6)
    return q


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed increments (0 = rest of the schedule)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed increments before timing (tare first)")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed", help="precision of the headline leg")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--nx", type=int, default=32)
    ap.add_argument("--ny", type=int, default=32)
    ap.add_argument("--nz", type=int, default=16)
    ap.add_argument("--load-steps", type=int, default=60)
    ap.add_argument("--unload-steps", type=int, default=60)
    ap.add_argument("--grad-tol", type=float, default=1e-7)
    ap.add_argument("--line-search", type=int, default=2, choices=[0, 1, 2])
    ap.add_argument("--strain", type=float, default=0.30)
    ap.add_argument("--max-iterations", type=int, default=20000)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--save-frames", default="", help="write each leg's frames to <prefix>_<precision>.npy")
    ap.add_argument("--workload", choices=["single", "large", "uneven", "sweep"], default="single",
                    help="single: the 32x32x16 lattice; large: the 64x64x32 contact-heavy lattice (a timed "
                         "window of --steps increments); uneven / sweep: batched lattice sets")
    ap.add_argument("--lattices", type=int, default=0, help="batch size (default 256 uneven, 1024 sweep)")
    ap.add_argument("--oracle-sample", type=int, default=0, help="batched: lattices checked against the oracle")
    ap.add_argument("--shards", type=int, default=0,
                    help="batched: independent contexts the batch is split over (MMC_BATCH_SHARDS)")
    ap.add_argument("--host-lattices", action="store_true",
                    help="batched: build the lattices on the host and upload them (default: on the device)")
    ap.add_argument("--single-sample", type=int, default=0,
                    help="batched: lattices also run one by one on the FP64 single-lattice solver")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this solver has no multi-GPU path (pass --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import mmc

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.workload in ("uneven", "sweep"):
        run_batch(args, log)
        return
    if args.workload == "large":
        if (args.nx, args.ny, args.nz) == (32, 32, 16):
            args.nx, args.ny, args.nz = 64, 64, 32
        args.load_steps, args.unload_steps = 100, 100
    prec = mmc.MMC_FP64 if args.precision == "fp64" else mmc.MMC_MIXED
    cfg, name = workload(args)
    other = mmc.MMC_FP64 if prec == mmc.MMC_MIXED else mmc.MMC_MIXED
    sim, fr, warm, K, dt, evals = run_leg(cfg, prec, args.warmup, args.steps, log, keep_state=prec == mmc.MMC_FP64)
    N, B = sim.n_nodes, sim.n_beams
    total = 1 + args.load_steps + args.unload_steps
    its = int(fr["iterations"][warm:].sum())
    # kernel durations with HIP events on the solver stream, at the state the
    # timed region ended in: the four kernels of a device-loop trial (launches
    # inside graph replays carry no per-launch events, so each is launched on
    # its own, on the loop's scratch buffers)
    launches = 200
    loop_us = {k: round(sim.time_kernel(k, launches) * 1e3, 3) for k in ("dstate_g", "elem_g", "gather_g", "decide_g")}
    # a plain 16 MiB device copy: the known-size reference of the traffic
    # counters (profiles/), and the copy bandwidth this device reaches
    copy_ms = sim.time_kernel("copy", 50)
    gather_ms = loop_us["gather_g"] * 1e-3
    g_bytes = gather_loop_bytes(N, B, cfg.history)
    ach = g_bytes / (gather_ms * 1e-3) / 1e9
    traffic = GATHER_TRAFFIC_BYTES.get((args.nx, args.ny, args.nz, args.grad_tol, args.precision))
    roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "k_gather_g",
            "bytes_per_launch": g_bytes, "mean_launch_us": round(gather_ms * 1e3, 3), "launches_timed": launches,
            # the L-BFGS vector traffic alone (200 B per DOF at a full history,
            # without the beam-end record reads of the gradient assembly)
            "vector_bytes_per_launch": 6 * N * (5 + 2 * cfg.history) * 8,
            "frac_vector_bytes": round(6 * N * (5 + 2 * cfg.history) * 8 / (gather_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "loop_kernel_us": loop_us, "trial_kernel_us": round(sum(loop_us.values()), 3),
            "loop_shape": sim.loop_shape(),
            "copy_16MiB_GBs": round(2 * (1 << 24) / (copy_ms * 1e-3) / 1e9, 1)}
    value = K / dt
    out = {
        "metric": "load-steps/sec, mixed vs FP64 on 1 GPU, at <=1e-5 rel platen-force error",
        "value": round(value, 4), "unit": "load-steps/s", "n_gpus": 1, "steps": K, "warmup": warm,
        "ms_per_step": round(dt / K * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if prec == 0 else "f32-element/f64-reduction",
        "data": "synthetic (seeded cubic lattice)",
        "config": {"workload": name, "nodes": N, "beams": B, "precision": args.precision,
                   "increments": f"{warm} warmup + {K} timed of {total}"},
        "lbfgs_iterations": its,
        "objective_evaluations": int(evals),
        "us_per_iteration": round(dt / max(1, its) * 1e6, 2),
        "beam_evaluations_per_s": round(evals * B / dt, 1),
        "peak_force_n": float(fr["reaction_force"].max()),
        "all_converged": bool(np.all(fr["converged"] == 1)),
        "roofline": roof,
    }
    state = sim.warm_state
    sim.close()
    if args.save_frames:
        np.save(f"{args.save_frames}_{args.precision}.npy", fr)
    fp64_dt, fp64_its, fr2 = dt, its, None
    if not args.no_compare:
        sim2, fr2, _, K2, dt2, _ = run_leg(cfg, other, args.warmup, args.steps, log, keep_state=other == mmc.MMC_FP64)
        if other == mmc.MMC_FP64:
            state = sim2.warm_state
            fp64_dt, fp64_its = dt2, int(fr2["iterations"][warm:].sum())
        sim2.close()
        lbl = "fp64" if other == mmc.MMC_FP64 else "mixed"
        if args.save_frames:
            np.save(f"{args.save_frames}_{lbl}.npy", fr2)
        out[f"{lbl}_leg"] = {"value": round(K2 / dt2, 4), "ms_per_step": round(dt2 / K2 * 1e3, 3),
                             "lbfgs_iterations": int(fr2["iterations"][warm:].sum()),
                             "us_per_iteration": round(dt2 / max(1, int(fr2["iterations"][warm:].sum())) * 1e6, 2),
                             "all_converged": bool(np.all(fr2["converged"] == 1))}
        mixed_fr, fp64_fr = (fr, fr2) if prec == mmc.MMC_MIXED else (fr2, fr)
        out["speedup_mixed_vs_fp64"] = round((K / dt) / (K2 / dt2) if prec == mmc.MMC_MIXED
                                             else (K2 / dt2) / (K / dt), 4)
        fo, fixture = oracle_fixture(args) if args.workload == "single" else (None, None)
        out["qualifier"] = qualifier(mixed_fr, fp64_fr, fo, fixture)
    if not args.no_cpu and args.workload == "large":
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_window(cfg, args.cpu_budget, its, K)
    elif not args.no_cpu and state is not None:
        log("cpu baseline ...")
        fp64_frames = fr if prec == mmc.MMC_FP64 else fr2
        top = float(fp64_frames["platen_y"][warm])
        out["cpu_baseline"] = cpu_baseline_sample(cfg, state, top, args.cpu_budget, fp64_its, K, fp64_dt)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
