"""Writes the command-line goldens under tests/golden/cli/ from the reference's
own program compiled from its sources (oracle/_ref/mc_ref, built by
oracle/Makefile from /root/reference): for each case the text outputs as the
program wrote them, its stdout and exit code, and for each META3 cache a digest
(header and diagnostics records; for the sim case also the snapshots of the
first, peak-force and last frames).  Run from the repository root:  python tests/golden/make_cli_goldens.py

Cases (arguments as the reference's CLI takes them, MC/src/main.cpp:157-181):
  validate_quick   --mode validate --quick (the reference's CTest)
  sim_specimen     --mode sim on the 2x4x2 specimen under the coarse schedule
                   (10 + 7 increments to 50%, 500 iterations, contact 1e6,
                   --skip-stl)
  convergence      --mode convergence (coarse, medium and fine schedules of the
                   specimen)
  stl_specimen     --mode stl on the specimen at STL quality 12
"""
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "tests"))
import meta3  # noqa: E402

OUT = Path(__file__).resolve().parent / "cli"
SPECIMEN = ["--cells-x", "2", "--cells-y", "4", "--cells-z", "2", "--load-steps", "10", "--unload-steps", "7",
            "--max-strain", "0.5", "--max-iterations", "500", "--contact-stiffness", "1000000"]
CASES = {
    "validate_quick": ["--mode", "validate", "--quick"],
    "sim_specimen": ["--mode", "sim"] + SPECIMEN + ["--skip-stl"],
    "convergence": ["--mode", "convergence"],
    "stl_specimen": ["--mode", "stl", "--cells-x", "2", "--cells-y", "4", "--cells-z", "2", "--stl-quality", "12"],
}
# the mesh of stl_specimen is the one of tests/golden/stl/specimen_q12 (same
# lattice and quality); only the text files are kept here
TEXT = ("VALIDATION.txt", "compression_metrics.csv", "summary.json", "convergence.csv",
        "graded_bcc_energy_absorber.validation.txt")


def meta3_digest(path, snapshots):
    h, d, s = meta3.read(path)
    out = dict(header=np.frombuffer(h.tobytes(), np.uint8), diagnostics=d)
    if snapshots:
        peak = int(np.argmax(d["reaction_force"]))
        pick = sorted({0, peak, len(d) - 1})
        out.update(frames=np.array(pick), snapshots=s[pick])
    return out


def main():
    exe = ROOT / "oracle" / "_ref" / "mc_ref"
    if not exe.exists():
        raise SystemExit(f"{exe} missing: make -C oracle cli (with the reference present)")
    for name, args in CASES.items():
        dst = OUT / name
        if dst.exists():
            shutil.rmtree(dst)
        dst.mkdir(parents=True)
        with tempfile.TemporaryDirectory() as tmp:
            out = Path(tmp) / "out"
            p = subprocess.run([str(exe)] + args + ["--output", str(out)], capture_output=True, text=True,
                               timeout=1800)
            (dst / "stdout.txt").write_text(p.stdout)
            (dst / "returncode.txt").write_text(f"{p.returncode}\n")
            for f in TEXT:
                if (out / f).exists():
                    shutil.copy(out / f, dst / f)
            for m in sorted(out.glob("*.meta3")):
                np.savez_compressed(dst / f"{m.stem}_meta3.npz", **meta3_digest(m, name == "sim_specimen"))
        print(name, "rc", p.returncode, sorted(x.name for x in dst.iterdir()))


if __name__ == "__main__":
    main()
