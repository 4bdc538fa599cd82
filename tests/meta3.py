"""Reader of the META3 snapshot cache (the reference's saveCache layout,
MC/src/main.cpp:9-92): a 32-byte header, then per frame one diagnostics record
and one record per body."""
import numpy as np

HEADER = np.dtype([("magic", "S8"), ("version", "<u4"), ("frames", "<u4"), ("bodies", "<u4"), ("fps", "<u4"),
                   ("snapshot_bytes", "<u4"), ("diagnostics_bytes", "<u4")])
SNAPSHOT = np.dtype([("p", "<f8", 3), ("q", "<f8", 4), ("v", "<f8", 3), ("w", "<f8", 3), ("half", "<f8", 3),
                     ("material", "<i4"), ("id", "<i4"), ("kind", "<i4"), ("module_class", "<i4"), ("ring", "<i4"),
                     ("slot", "<i4")])
DIAG_NAMES = ("time", "platen_y", "displacement", "engineering_strain", "engineering_stress", "reaction_force",
              "loading_work", "recovered_work", "dissipated_work", "specific_energy_absorption", "elastic_energy",
              "plastic_dissipation", "contact_energy", "residual_rms", "residual_max", "max_penetration",
              "mean_plastic_strain", "max_plastic_strain", "mean_plastic_rotation", "max_plastic_rotation",
              "permanent_set")
DIAGNOSTICS = np.dtype([(n, "<f8") for n in DIAG_NAMES] + [
    (n, "<i4") for n in ("active_contacts", "plastic_beams", "plastic_hinges", "iterations")] + [
    ("converged", "<u4"), ("loading", "<u4")])


def read(path):
    """(header, diagnostics[frames], snapshots[frames, bodies])."""
    raw = open(path, "rb").read()
    h = np.frombuffer(raw[:HEADER.itemsize], HEADER)[0]
    assert h["snapshot_bytes"] == SNAPSHOT.itemsize and h["diagnostics_bytes"] == DIAGNOSTICS.itemsize
    F, N = int(h["frames"]), int(h["bodies"])
    rec = np.dtype([("d", DIAGNOSTICS), ("s", SNAPSHOT, (N,))])
    body = np.frombuffer(raw[HEADER.itemsize:], rec)
    assert len(body) == F and len(raw) == HEADER.itemsize + F * rec.itemsize
    return h, body["d"].copy(), body["s"].copy()
