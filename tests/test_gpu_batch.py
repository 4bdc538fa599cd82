"""Batched lattices (mmc_batch_*) against the single-lattice facade.

A batch concatenates its lattices in one device context; every lattice keeps
its own block partition, reduction order and control block, so each lattice's
frames must equal, bit for bit, those of mmc_sim_run on that lattice alone
(which the parity tests pin against the oracle).  Both device-loop engines:
the workgroup-per-lattice solver runs each lattice in its own workgroup, the
graph-replayed loop runs every lattice's blocks in every replay (and re-captures
over the running ones as the batch drains).  The driver runs rounds of
mmc_minimize_batch (every lattice waits for the slowest of the round) or, with
MMC_BATCH_PER_LATTICE=1 and the graph loop, advances every lattice through its
own schedule (mmc_batch_start / advance / finish: a lattice starts its next
solve while others are still in theirs); both must give the single runs'
frames.  With MMC_BATCH_SHARDS > 1 the batch is split over independent
contexts run by host threads of their own, which must change nothing
either."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _configs(mmc):
    a = mmc.cubic_config(4, 4, 3, load=12, unload=8)
    b = mmc.cubic_config(5, 4, 3, load=10, unload=6, sigma=1e-4, radius_mode=1, r_bot=3e-4, r_top=7e-4)
    c = mmc.cubic_config(3, 3, 3, load=8, unload=8, strain=0.4).replace(young=1.3e9, yield_=35e6, hardening=0.08 * 1.3e9)
    d = mmc.default_config().replace(cells_x=2, cells_y=4, cells_z=2, max_strain=0.5, load_steps=10,
                                     unload_steps=7, max_iterations=500, contact_stiffness=1e6)
    return [a, b, c, d]


def _same(fa, fb):
    assert len(fa) == len(fb)
    for name in fa.dtype.names:
        assert np.array_equal(fa[name], fb[name]), name


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_single_runs(gpu, precision, loop_engine):
    mmc = gpu
    _batch_vs_single(mmc, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_per_lattice_equals_single_runs(gpu, precision, monkeypatch):
    monkeypatch.setenv("MMC_DEVICE_LOOP", "graph")
    monkeypatch.setenv("MMC_BATCH_PER_LATTICE", "1")
    _batch_vs_single(gpu, precision)


def _batch_vs_single(mmc, precision):
    prec = mmc.MMC_FP64 if precision == "fp64" else mmc.MMC_MIXED
    cfgs = _configs(mmc)
    frames, status, err = mmc.BatchSimulator(cfgs, prec).run()
    assert status == [0] * len(cfgs), err
    for cfg, fb in zip(cfgs, frames):
        fs = mmc.Simulator(cfg, prec).run()
        print(precision, len(fs), "frames, peak", fs["reaction_force"].max())
        _same(fb, fs)


def test_batch_isolates_a_failed_lattice(gpu, loop_engine):
    """A lattice whose equilibrium fails stops with the facade's error text;
    the others finish unchanged."""
    mmc = gpu
    cfgs = _configs(mmc)[:2]
    bad = cfgs[0].replace(max_iterations=3, max_adaptive_refinements=1, relative_escape=0)
    frames, status, err = mmc.BatchSimulator([bad, cfgs[1]], mmc.MMC_FP64).run()
    assert status[0] == mmc.MMC_ERR_EQUILIBRIUM and status[1] == 0
    with pytest.raises(mmc.MMCError) as e:
        mmc.Simulator(bad).run()
    assert str(e.value) in err
    _same(frames[1], mmc.Simulator(cfgs[1]).run())


def test_batch_per_lattice_isolates_a_failed_lattice(gpu, monkeypatch):
    monkeypatch.setenv("MMC_DEVICE_LOOP", "graph")
    monkeypatch.setenv("MMC_BATCH_PER_LATTICE", "1")
    test_batch_isolates_a_failed_lattice(gpu, "graph")


@pytest.mark.parametrize("shards", [2, 3])
def test_sharded_batch_equals_single_runs(gpu, shards, monkeypatch):
    """The four lattices over two and three shards (contexts on their own
    streams and host threads), lattice b in shard b mod shards."""
    monkeypatch.setenv("MMC_BATCH_SHARDS", str(shards))
    _batch_vs_single(gpu, "fp64")


@pytest.mark.parametrize("split", ["1", "16"])
def test_batch_decision_split_equals_single_runs(gpu, split, monkeypatch):
    """The decision kernel fixed at one block per lattice (every row folded by
    the deciding block, no arrival counter) and at sixteen, over the four
    lattices of one batch (the tuner's choice per set size bypassed), against
    single runs at the tuner's shape: bit for bit, mixed policy."""
    mmc = gpu
    monkeypatch.setenv("MMC_DEVICE_LOOP", "graph")
    cfgs = _configs(mmc)
    monkeypatch.setenv("MMC_DECIDE_BLOCKS", split)
    frames, status, err = mmc.BatchSimulator(cfgs, mmc.MMC_MIXED).run()
    monkeypatch.delenv("MMC_DECIDE_BLOCKS")
    assert status == [0] * len(cfgs), err
    for cfg, fb in zip(cfgs, frames):
        _same(fb, mmc.Simulator(cfg, mmc.MMC_MIXED).run())
