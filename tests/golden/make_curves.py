"""Writes the oracle force-displacement curves under tests/golden/.

These are long FP64 runs of the oracle (the CPU restatement of the reference
solver, pinned against the reference by tests/test_oracle_pin.py), kept so that
GPU tests and the benchmark can compare whole curves without re-running it.
Run from the repository root, one curve per call (each is single-threaded):

  python tests/golden/make_curves.py <curve> [max frames]

Each accepted frame is written as it comes (<curve>_oracle_frames.npy), with
<curve>_oracle_status.json holding the frame count, the oracle's error text if
its schedule ended in an equilibrium failure, and the seconds spent.

Curves:
  cubic_8x8x4_tight      8x8x4 cubic workload, 60 + 60 increments to 30%,
                         gradient tolerance 1e-8 (rms and max), per-term Armijo
                         and its approximate-Wolfe completion (line_search 2),
                         relative escape off.  At 1e-7 the oracle's own
                         unloading frames 62-69 stand up to 5.8e-6 of the peak
                         force from its 1e-8 curve; 1e-8 lies within 5.3e-8 of
                         the 1e-9 curve (permanent set 1.3e-7 relative)
  cubic_32x32x16         32x32x16 cubic workload, 60 + 60 increments to 30%,
                         gradient tolerance 1e-6, the reference's Armijo test,
                         20,000 iterations per solve
  cubic_32x32x16_tight   the same in the parity form (tolerance 1e-7,
                         line_search 2): the benchmark's workload
  cubic_32x32x16_tight8  the same at tolerance 1e-8 (run with a frame limit:
                         the first increments, for the window test)
  large_16x16x8_tight    the contact-heavy family (2% node jitter, per-beam
                         radii 0.3..0.7 mm, 50% strain, 100 + 100) on 16x16x8
                         cells, parity form at gradient tolerance 3e-8 (the
                         lattice softens past 13% strain, where 1e-7 leaves a
                         band of ~1e-6 of the peak force)
  uneven_<i>_tight       lattice i of the 256-lattice uneven set, parity form at
                         gradient tolerance 1e-9
  sweep_<i>_tight        design i of the 1,024-design sweep, likewise
"""
import json
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
# gradient tolerance of the batched-set curves: their final, unloaded frames
# stand free of the upper platen, held only by gravity, and the permanent set
# is the height of that soft equilibrium: at 1e-8 the oracle's own set stands
# 3.6e-6 (relative) from its 1e-9 value on sweep design 0 (force 1.5e-8)
BATCH_TOL = 1e-9


def curve_config(name: str) -> mmc.Config:
    fixed = {
        "cubic_8x8x4_tight": lambda: mmc.tight(mmc.cubic_config(8, 8, 4), tol=1e-8),
        "cubic_32x32x16": lambda: mmc.cubic_config(32, 32, 16, max_iterations=20000),
        "cubic_32x32x16_tight": lambda: mmc.tight(mmc.cubic_config(32, 32, 16, max_iterations=20000)),
        "cubic_32x32x16_tight8": lambda: mmc.tight(mmc.cubic_config(32, 32, 16, max_iterations=40000), tol=1e-8),
        "large_16x16x8_tight": lambda: mmc.tight(mmc.large_config(16, 16, 8), tol=3e-8),
    }
    if name in fixed:
        return fixed[name]()
    kind, idx, *rest = name.split("_")
    cfg = mmc.batch_workload(kind, 0, indices=[int(idx)])[0]
    return mmc.tight(cfg, tol=BATCH_TOL) if rest == ["tight"] else cfg


def main():
    name = sys.argv[1]
    limit = int(sys.argv[2]) if len(sys.argv) > 2 else None
    cfg = curve_config(name)
    t0 = time.time()
    status_path = OUT / f"{name}_oracle_status.json"
    frames_path = OUT / f"{name}_oracle_frames.npy"

    # a rerun replaces an earlier partial curve only once it holds more frames
    # (or is complete)
    have = 0
    if frames_path.exists() and status_path.exists() and not json.loads(status_path.read_text())["complete"]:
        have = len(np.load(frames_path))

    def save(frames, error=None, done=False):
        if len(frames) <= have and not done:
            return
        np.save(frames_path, frames, allow_pickle=False)
        status_path.write_text(json.dumps({
            "curve": name, "frames": int(len(frames)), "complete": done, "error": error,
            "frame_limit": limit, "seconds": round(time.time() - t0, 1)}, indent=1) + "\n")

    last = []

    def on_frame(frames):
        last[:] = [frames]
        save(frames)
        print(f"{name}: frame {len(frames)} strain {frames['engineering_strain'][-1]:.4f} "
              f"force {frames['reaction_force'][-1]:.6g} N, {int(frames['iterations'][-1])} iterations, "
              f"{time.time() - t0:.0f} s", flush=True)

    try:
        frames = po.Solver(cfg, "orc").run(on_frame=on_frame, limit=limit)
        save(frames, done=True)
    except RuntimeError as e:
        frames = last[0] if last else np.zeros(0, dtype=mmc.FRAME_DTYPE)
        save(frames, error=str(e), done=True)
    print(name, "frames", len(frames), "iterations", int(frames["iterations"].sum()),
          f"{time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
