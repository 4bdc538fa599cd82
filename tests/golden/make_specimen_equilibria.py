"""Writes tests/golden/bcc_222_equilibria_oracle.npz: the reference's 2x4x2
graded BCC specimen in the parity form (gradient tolerance 1e-9, line_search 2,
80,000 iterations per solve; at 1e-8 the oracle's own tare stands 1.7e-6 from
its 1e-9 value), the tare and the first two load increments of its
coarse schedule (0, 5% and 10% strain), each solved by the oracle from the
committed state and committed.  Arrays: strain, top, reaction, energy,
iterations, converged.  Run from the repository root (about two minutes):
  python tests/golden/make_specimen_equilibria.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "falling-dice_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
import mmc  # noqa: E402
import py_oracle as po  # noqa: E402

STRAINS = (0.0, 0.05, 0.10)


def config():
    return mmc.tight(mmc.specimen_config().replace(max_iterations=80000), tol=1e-9)


def main():
    o = po.Solver(config(), "orc")
    rows = []
    for s in STRAINS:
        top = o.top0 - s * o.height
        st, fin, _ = o.minimize(top)
        o.commit()
        rows.append((s, top, fin.reaction, fin.energy, st.iterations, st.converged))
        print(rows[-1], flush=True)
    cols = list(zip(*rows))
    np.savez(Path(__file__).resolve().parent / "bcc_222_equilibria_oracle.npz", strain=np.array(cols[0]),
             top=np.array(cols[1]), reaction=np.array(cols[2]), energy=np.array(cols[3]),
             iterations=np.array(cols[4]), converged=np.array(cols[5]))


if __name__ == "__main__":
    main()
