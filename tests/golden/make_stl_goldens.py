"""Writes the STL goldens under tests/golden/stl/ from the reference's own
Simulator::exportSTL (oracle/_ref/libmmcref.so, the reference compiled from its
sources by oracle/Makefile): per case the validation report the reference
writes next to the STL, and a digest of the mesh: triangle count, the header,
the sorted set of distinct vertices (float32, mm, as stored) and the enclosed
volume.  Run from the repository root:  python tests/golden/make_stl_goldens.py

Cases:
  specimen_q12   the 2x4x2 graded BCC specimen, radial segments 12 (voxel 0.5 mm)
  cubic_443_q10  a 4x4x3 cubic lattice (cell 5 mm, r 0.4..0.6 mm) injected into
                 the reference, radial segments 10 (voxel 0.6 mm)
"""
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "falling-dice_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))
import mmc  # noqa: E402
import py_oracle as po  # noqa: E402
import stlfile  # noqa: E402

OUT = Path(__file__).resolve().parent / "stl"


def cases():
    spec = mmc.specimen_config()
    cub = mmc.cubic_config(4, 4, 3)
    return {"specimen_q12": (spec, None, 12),
            "cubic_443_q10": (cub, po.Solver(cub, "orc").lattice(), 10)}


def main():
    if not po.available("ref"):
        raise SystemExit("oracle/_ref/libmmcref.so missing: make -C oracle with the reference present")
    OUT.mkdir(parents=True, exist_ok=True)
    for name, (cfg, lat, seg) in cases().items():
        r = po.Solver(cfg, "ref", lattice=lat)
        with tempfile.TemporaryDirectory() as tmp:
            path = Path(tmp) / "lattice.stl"
            r.export_stl(path, seg)
            d = stlfile.digest(path)
            (OUT / f"{name}.validation.txt").write_text((Path(tmp) / "lattice.validation.txt").read_text())
        np.savez_compressed(OUT / f"{name}.npz", **d)
        print(name, "triangles", int(d["triangles"]), "vertices", len(d["vertices"]), "volume mm3",
              float(d["volume_mm3"]))


if __name__ == "__main__":
    main()
