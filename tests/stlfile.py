"""Reader of binary STL files (80-byte header, uint32 count, 50-byte records:
normal, three vertices as float32, uint16 attribute) and the digest the STL
goldens keep: triangle count, header, the sorted distinct vertices and the
enclosed volume (mm^3, from the float32 vertices)."""
import numpy as np

RECORD = np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("attr", "<u2")])


def read(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[80:84], "<u4")[0])
    rec = np.frombuffer(raw[84:], RECORD)
    assert len(rec) == n and len(raw) == 84 + 50 * n
    return raw[:80], rec


def digest(path):
    header, rec = read(path)
    v = rec["v"].astype(np.float64)
    vol = float(np.sum(np.einsum("ij,ij->i", v[:, 0], np.cross(v[:, 1], v[:, 2]))) / 6.0)
    verts = np.unique(rec["v"].reshape(-1, 3), axis=0)
    return dict(triangles=np.array(len(rec)), header=np.frombuffer(header, np.uint8), vertices=verts,
                volume_mm3=np.array(vol))
