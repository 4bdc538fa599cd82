"""The cell triangulation of the STL export (CPU): the case table that
libmmc.so generates (mmc_stl_case_table) must give a closed, consistently
oriented surface on any field.  Pinned on random fields, by combinatorics only
(vertices are identified by the grid edge they sit on): every grid edge that
changes sign is used, every mesh edge is shared by exactly two triangles, once
in each direction, and no triangle repeats an edge.  The table is this build's
own (face-consistent contours, fanned loops); the reference's classic table is
not needed for a closed mesh."""
import numpy as np
import pytest


def _corner(c):
    return (c & 1, (c >> 1) & 1, (c >> 2) & 1)


def _edge_corners(e):
    a, m = e >> 2, e & 3
    lows = [c for c in range(8) if not (c >> a) & 1]
    c0 = lows[m]
    return c0, c0 | (1 << a), a


def test_case_table_shape(mmc):
    counts, tris = mmc.stl_case_table()
    assert counts[0] == 0 and counts[255] == 0
    assert counts.max() <= 5
    for cs in range(256):
        used = {int(e) for t in tris[cs][: counts[cs]] for e in t}
        crossing = set()
        for e in range(12):
            c0, c1, _ = _edge_corners(e)
            if ((cs >> c0) & 1) != ((cs >> c1) & 1):
                crossing.add(e)
        assert used == crossing, cs


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_case_table_closes_the_surface(mmc, seed):
    counts, tris = mmc.stl_case_table()
    n = 14
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(n, n, n))
    if seed % 2 == 0:  # smooth blobs as well as white noise
        g = np.linspace(-1, 1, n)
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        f = np.minimum(np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 0.6, np.sqrt((X - 0.5) ** 2 + Y ** 2) - 0.25) + 0.2 * f
    f[0, :, :] = f[-1, :, :] = f[:, 0, :] = f[:, -1, :] = f[:, :, 0] = f[:, :, -1] = 1.0  # closed inside the grid
    directed = {}
    ntri = 0
    for i in range(n - 1):
        for j in range(n - 1):
            for k in range(n - 1):
                cs = 0
                for c in range(8):
                    dx, dy, dz = _corner(c)
                    if f[i + dx, j + dy, k + dz] < 0.0:
                        cs |= 1 << c
                for t in tris[cs][: counts[cs]]:
                    keys = []
                    for e in t:
                        c0, _, a = _edge_corners(int(e))
                        dx, dy, dz = _corner(c0)
                        keys.append((a, i + dx, j + dy, k + dz))
                    assert len(set(keys)) == 3
                    ntri += 1
                    for u in range(3):
                        d = (keys[u], keys[(u + 1) % 3])
                        directed[d] = directed.get(d, 0) + 1
    assert ntri > 50
    for (p, q), cnt in directed.items():
        assert cnt == 1, (p, q)
        assert directed.get((q, p)) == 1, (p, q)
