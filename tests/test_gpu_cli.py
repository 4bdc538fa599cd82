"""The command-line program on the GPU against the goldens of the reference's
own program (tests/golden/cli/, written by tests/golden/make_cli_goldens.py
from oracle/_ref/mc_ref, the reference compiled from its sources).

Two programs run each case:
  * oracle/_ref/mc_dropin — the reference's MC/src/main.cpp and render.cpp
    compiled unchanged against the reference's headers and linked against
    libmmc.so in place of MC/src/sim.cpp: the source-level drop-in;
  * falling-dice_amd/metamaterial_compression — this package's own CLI, with
    the same options, outputs and exit codes (plus additive options).

What must agree exactly: the files written, every text key and its order, the
CSV header and row count, the schedule columns (frame, time, loading, platen,
displacement, strain), the lattice counts, mass and area, every PASS/FAIL flag,
the META3 header, body count, ids, kinds, classes, rings and slots, and the
exit code.  The numbers of an equilibrium run agree only to the band that the
reference's loose default tolerance leaves (rms 1.2e-3 / max 1.5e-2 or the
relative escape; a run of the same algorithm that sums in another order stops
elsewhere in it): force-like columns to 20% of the curve's peak, works to 12%
(dissipated work, a small difference of two works, 50%), permanent set 8%,
positions in the META3 snapshots per frame (POSITION_BANDS_M).  Those bands
are the spread measured over seven solvers of the specimen that differ only
in summation order (the GPU graph loop with three layouts of its reductions
and gradient records, FP64 and mixed, the block solver FP64 and mixed, the
host-driven loop): platen force 8.3-15.8% of the peak, loading work
4.5-8.8%, dissipated work 22-39%, permanent set 3.2-5.8%; snapshot
positions 5e-13 m at the tare (every solver starts from the same state and
the tare solve is short), 0.15 mm at the peak force, and 0.42 mm at the last
frame, the unloaded specimen standing free of the platen, whose shape is the
plastic state accumulated over every increment of the run.  The
finite-difference gradient check's relative
error (about 1e-3: the truncation error of the central differences at the
run's final state, which lies anywhere in the band) to 50% (measured 15%
apart); its PASS flag exactly.  The equilibria themselves are pinned at 1e-6
in tests/test_gpu_curves.py."""
import json
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import meta3
import stlfile
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
PROGRAMS = {"dropin": ROOT / "oracle" / "_ref" / "mc_dropin",
            "cli": ROOT / "falling-dice_amd" / "metamaterial_compression"}
CLI = GOLDEN / "cli"
# META3 snapshot position bands (m) at the tare, peak-force and last frames:
# 1e-9 where every solver agrees to round-off, elsewhere about 1.5x the
# largest difference measured between solvers that differ only in summation
# order (0.15 mm at the peak, 0.42 mm unloaded)
POSITION_BANDS_M = (1e-9, 3e-4, 6e-4)
SPECIMEN = ["--cells-x", "2", "--cells-y", "4", "--cells-z", "2", "--load-steps", "10", "--unload-steps", "7",
            "--max-strain", "0.5", "--max-iterations", "500", "--contact-stiffness", "1000000"]
CASES = {
    "validate_quick": ["--mode", "validate", "--quick"],
    "sim_specimen": ["--mode", "sim"] + SPECIMEN + ["--skip-stl"],
    "convergence": ["--mode", "convergence"],
    "stl_specimen": ["--mode", "stl", "--cells-x", "2", "--cells-y", "4", "--cells-z", "2", "--stl-quality", "12"],
}


def run_case(program, case, tmp_path):
    exe = PROGRAMS[program]
    if not exe.exists():
        pytest.fail(f"{exe} missing (__graft_entry__.build())")
    out = tmp_path / "out"
    print(f"running {program} {case}", flush=True)
    p = subprocess.run([str(exe)] + CASES[case] + ["--output", str(out)], capture_output=True, text=True,
                       timeout=600)
    want_rc = int((CLI / case / "returncode.txt").read_text())
    assert p.returncode == want_rc, p.stderr[-2000:]
    return out, p.stdout


def parse_validation(text):
    rows = []
    for line in text.splitlines():
        if "=" in line and not line.startswith("="):
            key, rest = line.split("=", 1)
            parts = rest.split()
            rows.append((key, parts[0], parts[1] if len(parts) > 1 else None))
    return rows


def check_validation(got_text, want_text, run):
    got, want = parse_validation(got_text), parse_validation(want_text)
    assert [r[0] for r in got] == [r[0] for r in want]
    assert [r[2] for r in got] == [r[2] for r in want], "PASS/FAIL/WARN flags"
    bands = {"peak_force_n": 0.20, "loading_work_j": 0.12, "recovered_work_j": 0.12, "dissipated_work_j": 0.50,
             "specific_energy_absorption_j_per_kg": 0.50, "permanent_set": 0.08,
             "max_contact_penetration_mm": 0.5, "finite_difference_gradient_relative_error": 0.50}
    for (k, g, _), (_, w, _) in zip(got, want):
        if k in ("nodes", "beams", "mass_kg", "reference_area_m2", "converged_step_fraction"):
            assert g == w, k
        elif k == "rigid_motion_objectivity_error":
            assert float(g) < 1e-9
        elif k in bands:
            assert abs(float(g) - float(w)) <= bands[k] * abs(float(w)), (k, g, w)
    return dict((k, v) for k, v, _ in got)


@pytest.mark.parametrize("program", ["dropin", "cli"])
def test_validate_quick_matches_reference(gpu, program, tmp_path):
    """--mode validate --quick, the reference's CTest (MC/CMakeLists.txt:15-20)."""
    out, stdout = run_case(program, "validate_quick", tmp_path)
    got = (out / "VALIDATION.txt").read_text()
    want = (CLI / "validate_quick" / "VALIDATION.txt").read_text()
    vals = check_validation(got, want, program)
    print(program, vals)
    assert sorted(p.name for p in out.iterdir()) == ["VALIDATION.txt"]


def _csv(path):
    lines = path.read_text().splitlines()
    header = lines[0].split(",")
    rows = [line.split(",") for line in lines[1:]]
    return header, rows


@pytest.mark.parametrize("program", ["dropin", "cli"])
def test_sim_specimen_outputs_match_reference(gpu, program, tmp_path):
    """--mode sim on the 2x4x2 specimen: the metrics CSV, summary.json,
    VALIDATION.txt, the META3 cache and the per-step stdout lines."""
    out, stdout = run_case(program, "sim_specimen", tmp_path)
    gold = CLI / "sim_specimen"
    assert sorted(p.name for p in out.iterdir()) == sorted(
        ["VALIDATION.txt", "compression_metrics.csv", "summary.json", "compression.meta3"])
    # metrics CSV
    gh, gr = _csv(out / "compression_metrics.csv")
    wh, wr = _csv(gold / "compression_metrics.csv")
    assert gh == wh and len(gr) == len(wr)
    col = {n: i for i, n in enumerate(wh)}
    for name in ("frame", "time", "loading", "platen_y_m", "displacement_m", "engineering_strain", "converged"):
        assert [r[col[name]] for r in gr] == [r[col[name]] for r in wr], name
    f_got = np.array([float(r[col["reaction_force_n"]]) for r in gr])
    f_want = np.array([float(r[col["reaction_force_n"]]) for r in wr])
    force_err = float(np.max(np.abs(f_got - f_want)) / f_want.max())
    print(program, "csv force error (of peak)", force_err)
    assert force_err <= 0.20
    # summary.json
    gs, ws = json.loads((out / "summary.json").read_text()), json.loads((gold / "summary.json").read_text())
    assert list(gs) == list(ws)
    for k in ("solver", "material", "cells", "frames", "mass_kg", "unconverged_steps"):
        assert gs[k] == ws[k], k
    for k, band in (("peak_force_n", 0.20), ("loading_work_j", 0.12), ("recovered_work_j", 0.12),
                    ("dissipated_work_j", 0.50), ("permanent_set", 0.08)):
        assert abs(gs[k] - ws[k]) <= band * abs(ws[k]), (k, gs[k], ws[k])
    # VALIDATION.txt
    check_validation((out / "VALIDATION.txt").read_text(), (gold / "VALIDATION.txt").read_text(), program)
    # META3 cache
    h, d, s = meta3.read(out / "compression.meta3")
    g = np.load(gold / "compression_meta3.npz")
    assert h.tobytes() == g["header"].tobytes()
    wd = g["diagnostics"]
    for name in ("time", "platen_y", "displacement", "engineering_strain", "loading", "converged"):
        assert np.array_equal(d[name], wd[name]), name
    picked = s[g["frames"]]
    ws_ = g["snapshots"]
    for name in ("id", "kind", "module_class", "ring", "slot"):
        assert np.array_equal(picked[name], ws_[name]), name
    per_frame = np.max(np.abs(picked["p"] - ws_["p"]), axis=(1, 2))
    print(program, "META3 max position difference (m) per snapshot frame", dict(zip(g["frames"].tolist(),
                                                                                   per_frame.tolist())))
    # snapshots at the tare, the peak force and the last (unloaded) frame
    assert g["frames"].tolist() == [0, 10, 17]
    for frame, dp, band in zip(g["frames"].tolist(), per_frame.tolist(), POSITION_BANDS_M):
        assert dp <= band, (frame, dp, band)
    # stdout: one [equilibrium] line per accepted step, then [final]
    steps = [ln for ln in stdout.splitlines() if ln.startswith("[equilibrium]")]
    want_steps = [ln for ln in (gold / "stdout.txt").read_text().splitlines() if ln.startswith("[equilibrium]")]
    assert len(steps) == len(want_steps)
    pat = re.compile(r"^\[equilibrium\] +\d+ (load  |unload) strain=\d\.\d{4} force=\d+\.\d{3} N "
                     r"residual=\d\.\d{2}e[-+]\d{2} pen=\d+\.\d{3} mm iter=\d+ \[(OK|REFINED)\]$")
    assert all(pat.match(ln) for ln in steps), steps[:3]
    assert [ln.split("strain=")[1].split()[0] for ln in steps] == [
        ln.split("strain=")[1].split()[0] for ln in want_steps]
    final = [ln for ln in stdout.splitlines() if ln.startswith("[final]")]
    want_final = [ln for ln in (gold / "stdout.txt").read_text().splitlines() if ln.startswith("[final]")]
    assert len(final) == 1 and [t.split("=")[0] for t in final[0].split()] == [
        t.split("=")[0] for t in want_final[0].split()]


@pytest.mark.parametrize("program", ["dropin", "cli"])
def test_convergence_study_matches_reference(gpu, program, tmp_path):
    """--mode convergence: the coarse, medium and fine schedules of the
    specimen, convergence.csv and one META3 cache per case."""
    out, stdout = run_case(program, "convergence", tmp_path)
    gold = CLI / "convergence"
    gh, gr = _csv(out / "convergence.csv")
    wh, wr = _csv(gold / "convergence.csv")
    assert gh == wh and len(gr) == len(wr) == 3
    for r, w in zip(gr, wr):
        assert r[:4] == w[:4]  # case, load_steps, unload_steps, max_iterations
        assert r[-1] == w[-1]  # unconverged steps
        assert abs(float(r[4]) - float(w[4])) <= 0.20 * float(w[4])  # peak force
        assert abs(float(r[5]) - float(w[5])) <= 0.12 * float(w[5])  # loading work
    for case in ("coarse", "medium", "fine"):
        h, d, s = meta3.read(out / f"convergence_{case}.meta3")
        g = np.load(gold / f"convergence_{case}_meta3.npz")
        assert h.tobytes() == g["header"].tobytes(), case
        assert np.array_equal(d["engineering_strain"], g["diagnostics"]["engineering_strain"]), case


@pytest.mark.parametrize("program", ["dropin", "cli"])
def test_stl_mode_matches_reference(gpu, program, tmp_path):
    """--mode stl on the specimen: the binary STL and its validation report
    (the mesh against tests/golden/stl/specimen_q12, the same lattice and
    quality: shared grid-edge vertices, watertight, volume within 0.5%; see
    tests/test_gpu_stl.py) and VALIDATION.txt."""
    out, stdout = run_case(program, "stl_specimen", tmp_path)
    gold = CLI / "stl_specimen"
    assert sorted(p.name for p in out.iterdir()) == sorted(
        ["VALIDATION.txt", "graded_bcc_energy_absorber.stl", "graded_bcc_energy_absorber.validation.txt"])
    rep = dict(ln.split("=", 1) for ln in (out / "graded_bcc_energy_absorber.validation.txt").read_text().splitlines()
               if "=" in ln)
    want = dict(ln.split("=", 1) for ln in (gold / "graded_bcc_energy_absorber.validation.txt").read_text().splitlines()
                if "=" in ln)
    assert list(rep) == list(want)
    assert rep["voxel_mm"] == want["voxel_mm"] and rep["grid"] == want["grid"] and rep["watertight"] == "PASS"
    d = stlfile.digest(out / "graded_bcc_energy_absorber.stl")
    g = np.load(GOLDEN / "stl" / "specimen_q12.npz")
    ours = {tuple(v) for v in d["vertices"].tolist()}
    theirs = {tuple(v) for v in g["vertices"].tolist()}
    print(program, "vertices", len(ours), len(theirs), "only ours", len(ours - theirs), "only reference",
          len(theirs - ours))
    assert len(ours - theirs) <= int(rep["repaired_boundary_loops"])
    assert len(theirs - ours) <= int(want["repaired_boundary_loops"])
    assert abs(float(d["volume_mm3"]) - float(g["volume_mm3"])) <= 5e-3 * float(g["volume_mm3"])
    check_validation((out / "VALIDATION.txt").read_text(), (gold / "VALIDATION.txt").read_text(), program)
