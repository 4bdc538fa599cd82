"""Whole force-displacement curves of the GPU path against committed oracle and
reference curves, at the parity tolerances of the solver's specification:
platen force within 1e-6 of the peak force on every frame for the FP64 path and
1e-5 for the mixed path, loading/dissipated work and permanent set likewise.

The curves run with the parity form of each workload (per-term Armijo test
with its approximate-Wolfe completion, relative escape off; mmc.tight) at a
gradient tolerance tight enough that the band it leaves lies inside the bound
(1e-8 for the cubic curves, 3e-8 for the contact-heavy family, 1e-9 for the
batched sets and the soft BCC specimen; each test states how it was chosen,
mostly from the oracle's own curve at a ten times tighter tolerance): at the
workload's own tolerance
(1e-6, the reference's total-energy Armijo test) two correct solvers stop
anywhere inside the band that tolerance leaves, ~1e-4 of the peak force on the
8x8x4 curve, so only a tighter solve can pin 1e-6.  Fixtures:
tests/golden/*_oracle_frames.npy (the oracle, tests/golden/make_curves.py) and
tests/golden/bcc_222_run_ref.npy (the reference compiled from its own sources,
tests/golden/make_golden.py)."""
import json

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

# parity bounds, relative to the oracle curve's peak platen force (and for the
# works and the permanent set, to the oracle's final values)
BOUND = {"fp64": 1e-6, "mixed": 1e-5}


def curve_errors(fg, fo):
    peak = max(float(np.max(fo["reaction_force"])), 1e-12)
    per_frame = np.abs(fg["reaction_force"] - fo["reaction_force"]) / peak
    out = {"force": float(per_frame.max()), "worst_frame": int(per_frame.argmax())}
    for k in ("loading_work", "recovered_work", "dissipated_work"):
        ref = max(abs(float(fo[k][-1])), 1e-12)
        out[k] = float(np.max(np.abs(fg[k] - fo[k]))) / ref
    # the permanent set of the run: the final frame's, relative
    out["permanent_set"] = abs(float(fg["permanent_set"][-1] - fo["permanent_set"][-1])) / max(
        abs(float(fo["permanent_set"][-1])), 1e-12)
    out["strain"] = float(np.max(np.abs(fg["engineering_strain"] - fo["engineering_strain"])))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config1_tight_curve_matches_oracle(gpu, precision):
    """8x8x4 cubic lattice, 30% strain in 60 + 60 increments (with the oracle's
    bisections), every frame against the oracle's curve, both at gradient
    tolerance 1e-8: at 1e-7 the oracle's own unloading frames (62-69) stand up
    to 5.8e-6 of the peak force from its 1e-8 curve, wider than the FP64
    bound; its 1e-8 curve is within 5.3e-8 of its 1e-9 curve."""
    mmc = gpu
    cfg = mmc.tight(mmc.cubic_config(8, 8, 4), tol=1e-8)
    fo = np.load(GOLDEN / "cubic_8x8x4_tight_oracle_frames.npy")
    prec = mmc.MMC_FP64 if precision == "fp64" else mmc.MMC_MIXED
    fg = mmc.Simulator(cfg, prec).run()
    assert len(fg) == len(fo), (len(fg), len(fo))
    assert np.all(fg["converged"] == 1)
    assert np.array_equal(fg["loading"], fo["loading"])
    err = curve_errors(fg, fo)
    print(precision, "config1 tight errors", err, "iterations", int(fg["iterations"].sum()),
          int(fo["iterations"].sum()))
    b = BOUND[precision]
    assert err["strain"] <= 1e-15
    assert err["force"] <= b
    assert err["loading_work"] <= b and err["recovered_work"] <= b and err["dissipated_work"] <= b
    assert err["permanent_set"] <= b
    assert np.all(fg["residual_max"] <= 1e-7) and np.all(fg["residual_rms"] <= 1e-7)


def test_published_specimen_run_matches_reference(gpu):
    """The 2x4x2 graded BCC specimen (MC/docs/VALIDATION.md:9-11) under the
    reference's coarse schedule (10 + 7 increments to 50% strain, self-contact,
    bisection; MC/src/main.cpp:213-221) against the curve of the reference
    compiled from its own sources.  These runs stop at the reference's own
    loose tolerance (rms 1.2e-3 and max 1.5e-2 in gradient units, or the
    relative escape at 1.5e-3), which leaves the specimen far from equilibrium:
    the reference's own force at 5% strain is 44% above the equilibrium the
    parity form finds (tests/golden/bcc_222_equilibria_oracle.npz).  Two
    solvers that sum in different orders stop at different points of that
    band.  The bands are the spread measured over six solvers that differ
    only in summation order (the graph loop before and after a change of its
    wave reductions, FP64 and mixed, the block solver FP64 and mixed, the
    host-driven loop), with a margin: platen force within 20% of the peak
    (measured 8.3-15.8%), loading work 12% (4.5-8.8%), dissipated work 50%
    (22-39%), permanent set 8% relative (3.2-5.8%); the schedule, the
    bisections and the convergence flags exactly."""
    mmc = gpu
    fr = np.load(GOLDEN / "bcc_222_run_ref.npy")
    fg = mmc.Simulator(mmc.specimen_config()).run()
    assert len(fg) == len(fr)
    assert np.array_equal(fg["converged"], fr["converged"])
    assert np.array_equal(fg["loading"], fr["loading"])
    assert float(np.max(np.abs(fg["engineering_strain"] - fr["engineering_strain"]))) <= 1e-15
    err = curve_errors(fg, fr)
    print("bcc 2x4x2 vs reference", err)
    assert err["force"] <= 0.20
    assert err["loading_work"] <= 0.12 and err["dissipated_work"] <= 0.50
    assert err["permanent_set"] <= 0.08


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_published_specimen_equilibria_match_oracle(gpu, po, precision):
    """The same specimen in the parity form against the oracle, which is the
    reference's algorithm (the oracle equals the reference bit for bit on this
    specimen, tests/test_oracle_pin.py): the tare and the first two load
    increments of its schedule (0, 5% and 10% strain), each solved from the
    committed state by both to gradient tolerance 1e-9 and committed (the
    oracle's side: tests/golden/bcc_222_equilibria_oracle.npz).  The
    soft, nearly incompressible BCC material (E 26 MPa, nu 0.485) is 77 times
    more compliant than the cubic workloads, so the same absolute gradient
    tolerance leaves a band 77 times wider: at 1e-8 the oracle's own tare
    stands 1.7e-6 of its force from its 1e-9 value; 1e-9 pins the 1e-6
    (GPU against oracle 1e-7 at the tare, 3e-8 and 2e-8 under load)."""
    mmc = gpu
    g = np.load(GOLDEN / "bcc_222_equilibria_oracle.npz")  # tests/golden/make_specimen_equilibria.py
    cfg = mmc.tight(mmc.specimen_config().replace(max_iterations=80000), tol=1e-9)
    lat = po.Solver(cfg, "orc").lattice()
    prec = mmc.MMC_FP64 if precision == "fp64" else mmc.MMC_MIXED
    ctx = mmc.Context(lat, cfg, prec)
    assert bool(np.all(g["converged"]))
    for k, top in enumerate(g["top"]):
        sg, fg, _ = ctx.minimize(float(top))
        err = abs(fg.reaction - g["reaction"][k]) / max(g["reaction"][k], 1.0)
        print(precision, "bcc 2x4x2 strain", g["strain"][k], "iterations", g["iterations"][k], sg.iterations,
              "force", g["reaction"][k], "error", err)
        assert sg.converged
        assert err <= BOUND[precision]
        assert abs(fg.energy - g["energy"][k]) <= BOUND[precision] * abs(g["energy"][k])
        ctx.commit()


def _advance_frames(sim, n):
    """The first n frames of a Simulator's schedule, increment by increment."""
    sim.begin()
    frames = []
    while sum(len(f) for f in frames) < n:
        f, done = sim.advance()
        frames.append(f)
        assert not done or sum(len(f) for f in frames) >= n
    return np.concatenate(frames)[:n]


def _fixture(name):
    path = GOLDEN / f"{name}_oracle_frames.npy"
    if not path.exists():
        pytest.fail(f"{path} missing (tests/golden/make_curves.py {name})")
    return np.load(path)


def _window_errors(fg, fo, n):
    """Per-frame force error over the first n frames, relative to the peak of
    the oracle's force over those frames."""
    peak = max(float(np.max(fo["reaction_force"][:n])), 1e-12)
    d = np.abs(fg["reaction_force"][:n] - fo["reaction_force"][:n]) / peak
    lw = abs(float(fg["loading_work"][n - 1] - fo["loading_work"][n - 1])) / max(float(fo["loading_work"][n - 1]), 1e-12)
    return float(d.max()), lw


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config2_window_matches_oracle(gpu, precision):
    """32x32x16 cubic lattice (18,513 nodes, 53,328 beams), the 30% schedule in
    parity form at gradient tolerance 1e-8: tare and the first 12 load
    increments (to 6% strain) against the oracle's curve.  At 1e-7 (the
    benchmark's tolerance) the two solvers agree to 1.1-1.4e-6 of the force
    over this window, depending on the GPU's summation order: the band 1e-7
    leaves is as wide as the FP64 bound."""
    mmc = gpu
    fo = _fixture("cubic_32x32x16_tight8")
    n = 13
    assert len(fo) >= n
    prec = mmc.MMC_FP64 if precision == "fp64" else mmc.MMC_MIXED
    cfg = mmc.tight(mmc.cubic_config(32, 32, 16, max_iterations=40000), tol=1e-8)
    fg = _advance_frames(mmc.Simulator(cfg, prec), n)
    assert np.array_equal(fg["engineering_strain"], fo["engineering_strain"][:n])
    assert np.all(fg["converged"] == 1)
    force, lw = _window_errors(fg, fo, n)
    print(precision, "32x32x16 window: force", force, "loading work", lw,
          "iterations", int(fg["iterations"].sum()), int(fo["iterations"][:n].sum()))
    assert force <= BOUND[precision] and lw <= BOUND[precision]


def _batched(mmc, names, precision):
    cfgs = []
    for name in names:
        kind, idx, _ = name.split("_")
        cfgs.append(mmc.tight(mmc.batch_workload(kind, 0, indices=[int(idx)])[0], tol=1e-9))
    prec = mmc.MMC_FP64 if precision == "fp64" else mmc.MMC_MIXED
    frames, status, err = mmc.BatchSimulator(cfgs, prec).run()
    assert status == [0] * len(cfgs), err
    for name, fg in zip(names, frames):
        fo = _fixture(name)
        b = BOUND[precision]
        if not json.loads((GOLDEN / f"{name}_oracle_status.json").read_text())["complete"]:
            # a curve still being written (tests/golden/make_curves.py): the
            # frames it holds, force and loading work over that window
            n = len(fo)
            assert n >= 40 and len(fg) > n, (name, len(fg), n)
            force, lw = _window_errors(fg, fo, n)
            print(precision, name, "cells", cfgs[names.index(name)].cells_x, "window", n, "force", force, "lw", lw)
            assert np.array_equal(fg["engineering_strain"][:n], fo["engineering_strain"])
            assert force <= b and lw <= b, name
            continue
        assert len(fg) == len(fo), (name, len(fg), len(fo))
        e = curve_errors(fg, fo)
        print(precision, name, "cells", cfgs[names.index(name)].cells_x, e)
        assert e["strain"] <= 1e-15
        assert e["force"] <= b, name
        assert e["loading_work"] <= b and e["dissipated_work"] <= b, name
        assert e["permanent_set"] <= b, name


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config5_sweep_designs_match_oracle(gpu, precision):
    """Four designs of the 1,024-design sweep (6^3 cells, per-design E, yield
    and hardening, 40% strain, 60 + 60), solved as one batch, each against the
    oracle's curve: force, loading and dissipated work (the sweep's energy-
    absorption figure) and permanent set.  Gradient tolerance 1e-9: the
    unloaded lattice stands free under gravity alone, and its height (the
    permanent set) is the softest equilibrium of the curve: at 1e-8 the
    oracle's own set stands 3.6e-6 from its 1e-9 value (design 0, force
    1.5e-8); at 1e-9 the GPU agrees with the oracle to 7e-9."""
    _batched(gpu, [f"sweep_{i}_tight" for i in (0, 1, 2, 3)], precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config4_uneven_lattices_match_oracle(gpu, precision):
    """Five lattices of the 256-lattice uneven set (5^3, 6^3, 6^3, 8^3 and 16^3
    cells, per-lattice radius grading, 30% strain, 60 + 60), solved as one
    batch of uneven sizes, each against the oracle's curve."""
    _batched(gpu, [f"uneven_{i}_tight" for i in (13, 5, 19, 4, 0)], precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config3_family_window_matches_oracle(gpu, precision):
    """The contact-heavy family (2% node jitter, per-beam radii 0.3..0.7 mm,
    50% strain in 100 + 100 increments) on 16x16x8 cells: the first 34 frames
    (to 16.5% strain: the force peaks at 13.5% and the lattice then softens)
    against the oracle's curve.  Gradient tolerance 3e-8: at 1e-7 the softening
    frames agree only to ~1e-6 of the peak force (measured 1.08e-6 FP64).  The
    window ends where the contact penetration reaches the reference's limit
    (0.12 mm): from 16.6% strain the oracle itself needs bisection there, and a
    solver that sums in another order may pass or fail the limit."""
    mmc = gpu
    fo = _fixture("large_16x16x8_tight")
    n = min(34, len(fo))
    assert n >= 30
    prec = mmc.MMC_FP64 if precision == "fp64" else mmc.MMC_MIXED
    fg = _advance_frames(mmc.Simulator(mmc.tight(mmc.large_config(16, 16, 8), tol=3e-8), prec), n)
    assert np.array_equal(fg["engineering_strain"], fo["engineering_strain"][:n])
    force, lw = _window_errors(fg, fo, n)
    print(precision, "16x16x8 jittered window: force", force, "loading work", lw)
    assert force <= BOUND[precision] and lw <= BOUND[precision]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config3_full_size_objective_matches_oracle(gpu, po, precision):
    """64x64x32 jittered lattice (139,425 nodes, 409,760 beams) at a seeded
    squashed, yielded state with heavy self-contact: one objective evaluation
    against the oracle's (about ten seconds on one host core).  FP64: energy and
    reaction 1e-12, gradient 1e-10 of its largest component; mixed: energy
    1e-7, gradient 2e-5."""
    from conftest import rel, seeded_z
    mmc = gpu
    cfg = mmc.large_config()
    o = po.Solver(cfg, "orc")
    lat = o.lattice()
    rng = np.random.default_rng(21)
    B = o.n_beams
    p = np.zeros((B, 7))
    p[:, 0] = rng.normal(0, 0.005, B)
    p[:, 1] = np.abs(rng.normal(0, 0.005, B))
    p[:, 2:5] = rng.normal(0, 0.01, (B, 3))
    p[:, 5:7] = np.abs(rng.normal(0, 0.005, (B, 2)))
    o.set_plastic(p.reshape(-1))
    lat.plastic = p.reshape(-1).copy()
    prec = mmc.MMC_FP64 if precision == "fp64" else mmc.MMC_MIXED
    ctx = mmc.Context(lat, cfg, prec)
    y = lat.x.reshape(-1, 3)[:, 1]
    # layers squashed to 15% of their spacing: ~3e5 sphere/capsule contacts
    z = seeded_z(o.n_nodes, 31, t_sigma=0.01, r_sigma=0.05, squash=0.85, y=y,
                 jr=cfg.joint_scale * cfg.max_radius, ts=cfg.cell_size)
    top = o.top0 - 0.85 * o.height
    ro, go = o.objective(z, top)
    rg, gg = ctx.objective(z, top)
    e_rel = abs(rg.energy - ro.energy) / abs(ro.energy)
    g_rel = rel(gg, go)
    print(precision, "64x64x32 contacts", ro.contacts, rg.contacts, "energy", e_rel, "grad", g_rel)
    assert ro.contacts > 100000 and rg.contacts == ro.contacts
    assert abs(rg.reaction - ro.reaction) <= 1e-12 * max(ro.reaction, 1.0)
    if precision == "fp64":
        assert e_rel <= 1e-12 and g_rel <= 1e-10
    else:
        assert e_rel <= 1e-7 and g_rel <= 2e-5
