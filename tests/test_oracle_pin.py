"""Pins the FP64 CPU oracle against the reference solver compiled from source.

The fixtures in tests/golden/ were written by tests/golden/make_golden.py from
oracle/_ref/libmmcref.so, i.e. the reference's own MC/src/sim.cpp.  The oracle
keeps the reference's operation order and hash-grid visiting order, so every
comparison here is exact (bit for bit), including the full load/unload runs
with self-contact and adaptive bisection.
"""
import numpy as np
import pytest

from conftest import GOLDEN


def bcc_222(mmc):
    return mmc.default_config().replace(cells_x=2, cells_y=4, cells_z=2, max_strain=0.5, load_steps=10,
                                        unload_steps=7, max_iterations=500, contact_stiffness=1e6)


def cubic_444(mmc):
    return mmc.cubic_config(4, 4, 4, load=4, unload=2, grad_tol=1e-3).replace(
        relative_escape=1, gradient_max_tolerance=1.5e-2, max_iterations=600)


def test_published_specimen_size(mmc, po):
    """2x4x2 graded BCC: 365 nodes, 456 beams, 6.413882 g (MC/docs/VALIDATION.md:9-11)."""
    s = po.Solver(bcc_222(mmc), "orc")
    assert (s.n_nodes, s.n_beams) == (365, 456)
    assert abs(s.mass - 0.006413882) < 5e-10
    assert abs(s.area - 0.0004) < 1e-15


def test_bcc_lattice_fixture(mmc, po):
    g = np.load(GOLDEN / "bcc_222_ref.npz")
    L = po.Solver(bcc_222(mmc), "orc").lattice()
    for k in ("x", "mass", "locked", "a", "b", "L0", "radius", "e0", "node_layer"):
        assert np.array_equal(getattr(L, k), g[k]), k
    assert np.array_equal(np.sort(L.connected), g["connected"])


@pytest.mark.parametrize("state", ["s0", "s1"])
def test_bcc_objective_fixture(mmc, po, state):
    g = np.load(GOLDEN / "bcc_222_ref.npz")
    s = po.Solver(bcc_222(mmc), "orc")
    out, grad, trial = s.objective(g[f"{state}_z"], float(g[f"{state}_top"][0]), trials=True)
    rec = np.array([out.energy, out.elastic, out.plastic, out.contact, out.reaction, out.max_pen, out.rms,
                    out.max, out.contacts])
    assert np.array_equal(rec, g[f"{state}_rec"])
    assert np.array_equal(grad, g[f"{state}_grad"])
    assert np.array_equal(trial, g[f"{state}_trial"])
    assert out.contacts > 0


def test_bcc_checks_fixture(mmc, po):
    g = np.load(GOLDEN / "bcc_222_ref.npz")
    s = po.Solver(bcc_222(mmc), "orc")
    ok, e = s.rigid_rotation_test()
    assert ok and e == g["objectivity"][0]
    ok, e = s.gradient_check()
    assert ok and e == g["fdcheck"][0]


def test_bcc_run_fixture(mmc, po):
    """10 + 7 increments to 50% strain with self-contact and bisection."""
    g = np.load(GOLDEN / "bcc_222_run_ref.npy")
    f = po.Solver(bcc_222(mmc), "orc").run()
    assert f.tobytes() == g.tobytes()


@pytest.mark.parametrize("state", ["random", "squash", "yielded"])
def test_cubic_objective_fixture(mmc, po, state):
    g = np.load(GOLDEN / "cubic_444_ref.npz")
    cfg = cubic_444(mmc)
    lat = po.Solver(cfg, "orc").lattice()
    s = po.Solver(cfg, "orc", lattice=lat)
    if f"{state}_plastic" in g:
        s.set_plastic(g[f"{state}_plastic"])
    out, grad, trial = s.objective(g[f"{state}_z"], float(g[f"{state}_top"][0]), trials=True)
    rec = np.array([out.energy, out.elastic, out.plastic, out.contact, out.reaction, out.max_pen, out.rms,
                    out.max, out.contacts])
    assert np.array_equal(rec, g[f"{state}_rec"])
    assert np.array_equal(grad, g[f"{state}_grad"])
    assert np.array_equal(trial, g[f"{state}_trial"])


def test_cubic_run_fixture(mmc, po):
    g = np.load(GOLDEN / "cubic_444_ref.npz")
    f = po.Solver(cubic_444(mmc), "orc").run()
    assert f.tobytes() == g["run_frames"].tobytes()


def test_cubic_curve_fixture_consistency():
    """The 8x8x4 curve used by the GPU parity test: converged everywhere, energy
    accounting signs as the reference's validation report requires."""
    f = np.load(GOLDEN / "cubic_8x8x4_oracle_frames.npy")
    assert len(f) == 121 and np.all(f["converged"] == 1)
    assert np.all(f["residual_rms"] <= 1e-6) and np.all(f["residual_max"] <= 1e-6)
    assert f["loading_work"][-1] + 1e-10 >= f["recovered_work"][-1]
    assert np.all(np.diff(f["loading_work"]) >= 0)


# ---- live comparison with the reference compiled from source (when built) ----
needs_ref = pytest.mark.skipif(not __import__("py_oracle").available("ref"),
                               reason="oracle/_ref not built (reference sources absent)")


@needs_ref
@pytest.mark.parametrize("cells,seed", [((3, 4, 3), 0x494E445553545259), ((2, 3, 2), 12345), ((3, 2, 2), 7)])
def test_live_reference_bcc(mmc, po, cells, seed):
    cfg = mmc.default_config().replace(cells_x=cells[0], cells_y=cells[1], cells_z=cells[2], seed=seed)
    o, r = po.Solver(cfg, "orc"), po.Solver(cfg, "ref")
    lo, lr = o.lattice(), r.lattice()
    for k in ("x", "mass", "locked", "a", "b", "L0", "radius", "area", "inertia", "e0"):
        assert np.array_equal(getattr(lo, k), getattr(lr, k)), k
    rng = np.random.default_rng(seed & 0xFFFF)
    z = rng.normal(0, 0.02, 6 * o.n_nodes)
    for top in (o.top0, o.top0 - 0.2 * o.height):
        a, ga, ta = o.objective(z, top, True)
        b, gb, tb = r.objective(z, top, True)
        assert a.energy == b.energy and a.contacts == b.contacts
        assert np.array_equal(ga, gb) and np.array_equal(ta, tb)


@needs_ref
def test_live_reference_minimize_commit(mmc, po):
    cfg = cubic_444(mmc)
    lat = po.Solver(cfg, "orc").lattice()
    o, r = po.Solver(cfg, "orc", lattice=lat), po.Solver(cfg, "ref", lattice=lat)
    for s in (0.0, 0.02, 0.04):
        so, fo, zo = o.minimize(o.top0 - s * o.height)
        sr, fr, zr = r.minimize(r.top0 - s * r.height)
        assert so.iterations == sr.iterations and so.backtracks == sr.backtracks
        assert np.array_equal(zo, zr) and fo.energy == fr.energy
        o.commit()
        r.commit()
        for a, b in zip(o.read_state(), r.read_state()):
            assert np.array_equal(a, b)


def test_capture_fixture(mmc, po):
    """Simulator::capture (MC/src/sim.cpp:215-310) at the fixture's committed
    pose and plastic state: with velocities against its previous pose and
    without (the first frame), every body bit for bit."""
    g = np.load(GOLDEN / "bcc_222_capture_ref.npz")
    s = po.Solver(bcc_222(mmc), "orc")
    s.set_pose(g["x"], g["q"])
    s.set_plastic(g["plastic"])
    top = float(g["top"][0])
    assert s.capture(top, g["prev_x"], g["prev_q"]).tobytes() == g["bodies"].tobytes()
    assert s.capture(top).tobytes() == g["bodies_first"].tobytes()


@needs_ref
@pytest.mark.parametrize("seed", [1, 2])
def test_live_reference_capture(mmc, po, seed):
    """capture after a committed minimisation (yielded struts, contact) against
    the reference's, velocities from the pose before it."""
    cfg = bcc_222(mmc)
    o, r = po.Solver(cfg, "orc"), po.Solver(cfg, "ref")
    top = o.top0 - 0.1 * seed * o.height
    px, pq, _ = o.read_state()
    for s in (o, r):
        s.minimize(top)
        s.commit()
    assert o.capture(top, px, pq).tobytes() == r.capture(top, px, pq).tobytes()
    assert o.capture(top).tobytes() == r.capture(top).tobytes()
