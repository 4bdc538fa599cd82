"""Simulator::capture (MC/src/sim.cpp:215-310) on the device, mmc_capture.

The oracle's capture is pinned bit for bit to the reference's
(tests/test_oracle_pin.py: a golden fixture and live runs) and by closed forms
(tests/test_oracle_math.py).  Here the device capture of committed GPU states
is compared with the oracle's capture of the same states, read back from the
device: the kernel computes without fused multiply-adds, so positions,
orientations, half extents, velocities and every label are bit for bit the
oracle's; angular velocities go through the device's atan2, which may differ
from the host's in the last place, and are held to 1e-14 of their largest
magnitude.  Batched runs hand every committed frame's bodies to an observer
(mmc_batch_set_frame_observer): each lattice's bodies equal those of its run
alone, bit for bit, and each frame's velocities are its joints' displacement
since the previous frame over the step time."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXACT = ("p", "q", "v", "half", "material", "id", "kind", "module_class", "ring", "slot")


def _check(dev, ref):
    assert len(dev) == len(ref)
    for name in EXACT:
        assert np.array_equal(dev[name], ref[name]), name
    wmax = float(np.abs(ref["w"]).max())
    assert float(np.abs(dev["w"] - ref["w"]).max()) <= 1e-14 * max(wmax, 1e-300)


def _configs(mmc):
    return {
        "bcc_222": mmc.default_config().replace(cells_x=2, cells_y=4, cells_z=2, max_iterations=400,
                                                contact_stiffness=1e6),
        "cubic_4x5x3_jitter": mmc.cubic_config(4, 5, 3, sigma=1e-4, radius_mode=1, r_bot=3e-4, r_top=7e-4),
    }


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["bcc_222", "cubic_4x5x3_jitter"])
def test_capture_matches_oracle(gpu, po, case, precision):
    """Three committed increments from a yielded state (all four strut material
    classes), each captured with velocities against the pose before it and in
    the first-frame form."""
    mmc = gpu
    cfg = _configs(mmc)[case]
    o = po.Solver(cfg, "orc")
    lat = o.lattice()
    rng = np.random.default_rng(3)
    p = np.zeros((o.n_beams, 7))
    p[:, 0] = rng.normal(0, 0.08, o.n_beams)
    p[:, 2:5] = rng.normal(0, 0.25, (o.n_beams, 3))
    lat.plastic = p.reshape(-1).copy()
    ctx = mmc.Context(lat, cfg, mmc.MMC_FP64 if precision == "fp64" else mmc.MMC_MIXED)
    for k, strain in enumerate((0.0, 0.01, 0.02)):
        top = o.top0 - strain * o.height
        x0, q0, _ = ctx.read_state()
        ctx.capture_mark()
        ctx.minimize(top)
        ctx.commit()
        x1, q1, p1 = ctx.read_state()
        o.set_pose(x1, q1)
        o.set_plastic(p1)
        dev = ctx.capture(top, velocities=True)
        ref = o.capture(top, x0, q0)
        _check(dev, ref)
        _check(ctx.capture(top), o.capture(top))
        if k == 2:
            assert set(np.unique(dev["material"][:o.n_beams])) == {0, 1, 2, 3}
            assert float(np.abs(dev["w"]).max()) > 0 and float(np.abs(dev["v"]).max()) > 0


def _run_with_bodies(mmc, cfgs, prec=0):
    bodies = {}
    frames, status, err = mmc.BatchSimulator(cfgs, prec).run(
        on_bodies=lambda lat, frame, b: bodies.__setitem__((lat, frame), b))
    assert status == [0] * len(cfgs), err
    return frames, bodies


def _batch_cfgs(mmc):
    return [mmc.default_config().replace(cells_x=2, cells_y=3, cells_z=2, max_strain=0.3, load_steps=4,
                                         unload_steps=2, max_iterations=400),
            mmc.cubic_config(3, 3, 3, load=4, unload=3, strain=0.2),
            mmc.cubic_config(4, 3, 3, load=3, unload=2, sigma=1e-4, radius_mode=1, r_bot=3e-4, r_top=7e-4)]


@pytest.mark.parametrize("driver", ["rounds", "per_lattice", "sharded"])
def test_batch_frames_carry_their_bodies(gpu, driver, monkeypatch):
    mmc = gpu
    if driver == "per_lattice":
        monkeypatch.setenv("MMC_DEVICE_LOOP", "graph")
        monkeypatch.setenv("MMC_BATCH_PER_LATTICE", "1")
    cfgs = _batch_cfgs(mmc)
    if driver == "sharded":
        monkeypatch.setenv("MMC_BATCH_SHARDS", "2")
    frames, bodies = _run_with_bodies(mmc, cfgs)
    monkeypatch.delenv("MMC_BATCH_SHARDS", raising=False)
    assert set(bodies) == {(b, f) for b, fr in enumerate(frames) for f in range(len(fr))}
    for b, cfg in enumerate(cfgs):
        alone, ab = _run_with_bodies(mmc, [cfg])
        assert len(alone[0]) == len(frames[b])
        inv_dt = 1.0 / max(1e-6, cfg.step_time)
        for f in range(len(frames[b])):
            got = bodies[(b, f)]
            assert got.tobytes() == ab[(0, f)].tobytes(), (b, f)
            n_beams = int(np.sum(got["kind"] == 0))
            joints = got[got["kind"] == 1]
            # the moving platen is the frame's platen
            assert got["p"][-6][1] == frames[b]["platen_y"][f] + 0.0030
            if f == 0:
                assert np.all(got["v"] == 0) and np.all(got["w"] == 0)
            else:
                prev = bodies[(b, f - 1)][n_beams:n_beams + len(joints)]
                assert np.array_equal(joints["v"], (joints["p"] - prev["p"]) * inv_dt), (b, f)
