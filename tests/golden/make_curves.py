"""Writes the oracle force-displacement curves under tests/golden/.

These are long FP64 runs of the oracle (the CPU restatement of the reference
solver, pinned against the reference by tests/test_oracle_pin.py), kept so that
GPU tests and the benchmark can compare whole curves without re-running it.
Run from the repository root, one curve per call (each is single-threaded):

  python tests/golden/make_curves.py cubic_8x8x4_tight
  python tests/golden/make_curves.py cubic_32x32x16
  python tests/golden/make_curves.py cubic_32x32x16_tight

Curves:
  cubic_8x8x4_tight      8x8x4 cubic workload, 60 + 60 increments, gradient
                         tolerance 1e-7 (rms and max), per-term Armijo
                         (line_search 1), relative escape off
  cubic_32x32x16         32x32x16 cubic workload, 60 + 60 increments, gradient
                         tolerance 1e-6, the reference's Armijo test
  cubic_32x32x16_tight   32x32x16 at tolerance 1e-7, per-term Armijo
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "falling-dice_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
import mmc  # noqa: E402
import py_oracle as po  # noqa: E402

OUT = Path(__file__).resolve().parent


def tight(cfg):
    return cfg.replace(line_search=1, gradient_rms_tolerance=1e-7, gradient_max_tolerance=1e-7, relative_escape=0)


CURVES = {
    "cubic_8x8x4_tight": lambda: tight(mmc.cubic_config(8, 8, 4)),
    "cubic_32x32x16": lambda: mmc.cubic_config(32, 32, 16),
    "cubic_32x32x16_tight": lambda: tight(mmc.cubic_config(32, 32, 16)),
}


def main():
    name = sys.argv[1]
    cfg = CURVES[name]()
    t0 = time.time()
    frames = po.Solver(cfg, "orc").run()
    np.save(OUT / f"{name}_oracle_frames.npy", frames)
    print(name, "frames", len(frames), "peak", frames["reaction_force"].max(),
          "iterations", int(frames["iterations"].sum()), f"{time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
