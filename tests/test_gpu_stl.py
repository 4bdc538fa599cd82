"""STL export on the GPU (mmc_export_stl, Simulator::exportSTL) against the
reference's own exportSTL (tests/golden/stl/, tests/golden/make_stl_goldens.py).

The signed-distance field, the grid and the snapped edge vertices are the
reference's; the cell triangulation is this build's (a face-consistent table),
so the two meshes share their vertices but not their triangles.  Both drop
zero-area slivers and close the holes they leave; the reference's table also
leaves holes at ambiguous cells, this one none.  What must agree:
  * the report's voxel size and grid exactly, and both watertight (no boundary
    and no non-manifold edge);
  * the vertex sets: every vertex on a grid edge is shared; each mesh may hold
    at most one vertex of its own per boundary loop its repair closes (the
    loop's centroid, summed in its own order and snapped to 1e-8 m);
  * the enclosed volume, within 0.5% (two triangulations of the same edge
    vertices differ inside each cell by a small fraction of a voxel's volume);
  * the 80-byte header."""
from pathlib import Path

import numpy as np
import pytest

import stlfile
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _report(text):
    return dict(line.split("=", 1) for line in text.splitlines() if "=" in line)


CASES = {"specimen_q12": 12, "cubic_443_q10": 10}


@pytest.mark.parametrize("case", list(CASES))
def test_stl_matches_reference(gpu, case, tmp_path):
    mmc = gpu
    cfg = mmc.specimen_config() if case.startswith("specimen") else mmc.cubic_config(4, 4, 3)
    sim = mmc.Simulator(cfg)
    path = tmp_path / "lattice.stl"
    rep = sim.export_stl(path, CASES[case])
    got = _report((tmp_path / "lattice.validation.txt").read_text())
    want = _report((GOLDEN / "stl" / f"{case}.validation.txt").read_text())
    print(case, "report", got, "reference", want)
    assert got["voxel_mm"] == want["voxel_mm"] and got["grid"] == want["grid"]
    assert got["watertight"] == "PASS" and got["boundary_edges"] == "0" and got["nonmanifold_edges"] == "0"
    # holes before the repair come from zero-area slivers (dropped in both);
    # the reference's ambiguous cells add more
    assert rep["watertight"] == 1 and rep["initial_boundary_edges"] <= int(want["initial_ambiguous_boundary_edges"])
    g = np.load(GOLDEN / "stl" / f"{case}.npz")
    d = stlfile.digest(path)
    out = Path(__file__).resolve().parents[1] / "gpurun_out"
    if out.is_dir():
        np.savez_compressed(out / f"stl_{case}.npz", **d)
    assert d["header"].tobytes() == g["header"].tobytes()
    ours = {tuple(v) for v in d["vertices"].tolist()}
    theirs = {tuple(v) for v in g["vertices"].tolist()}
    extra_ours, extra_theirs = ours - theirs, theirs - ours
    print(case, "vertices", len(ours), len(theirs), "only ours", len(extra_ours), "only reference",
          len(extra_theirs), "triangles", int(d["triangles"]), int(g["triangles"]))
    # vertices on grid edges are shared; the meshes differ only in the
    # centroids of the loops each repair closes (its own summation order)
    assert len(extra_ours) <= rep["repaired_loops"]
    assert len(extra_theirs) <= int(want["repaired_boundary_loops"])
    dv = abs(float(d["volume_mm3"]) - float(g["volume_mm3"])) / float(g["volume_mm3"])
    print(case, "volume mm3", float(d["volume_mm3"]), float(g["volume_mm3"]), "relative difference", dv)
    assert dv <= 5e-3
