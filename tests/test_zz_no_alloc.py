"""The allocation-free beam potential (mmc_beam_potential_noalloc) against the
allocating entry point (mmc_beam_potential), through the C-ABI.

The no-allocation variant takes device arrays and keeps its reduction
partials in a static device array; it shares the kernels of the allocating
entry point, whose partials live in the context's own buffer, so the two must
agree bit for bit.  Two calls in a row on different lattices (the second
smaller) check that nothing of the first call leaks into the second, and a
capture of the call into a HIP graph (global capture mode, under which an
allocation during capture is an error) replayed three times with eager calls
between the replays checks the graph-capture use."""
import numpy as np
import pytest

import hiprt

pytestmark = pytest.mark.gpu


def _case(mmc, po, cfg, seed):
    o = po.Solver(cfg, "orc")
    lat = o.lattice()
    ctx = mmc.Context(lat, cfg)
    rng = np.random.default_rng(seed)
    x = lat.x + rng.normal(0, 1e-4, lat.x.shape)
    q = lat.q.reshape(-1, 4) + rng.normal(0, 0.01, (o.n_nodes, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).reshape(-1)
    # device planes x[3][N], q[4][N] (w, x, y, z)
    xd = hiprt.DeviceArray(x.reshape(-1, 3).T)
    qd = hiprt.DeviceArray(q.reshape(-1, 4).T)
    return ctx, x, q, xd, qd


def _call(mmc, ctx, xd, qd, e, stream=None):
    rc = mmc.lib().mmc_beam_potential_noalloc(ctx.h, xd.ptr, qd.ptr, e.ptr, stream)
    if rc != 0:
        raise mmc.MMCError(mmc.lib().mmc_last_error(ctx.h).decode())


def test_noalloc_beam_potential_equals_allocating_entry(gpu, po):
    mmc = gpu
    big = _case(mmc, po, mmc.cubic_config(8, 8, 4), 4)
    small = _case(mmc, po, mmc.cubic_config(4, 4, 3), 5)
    e = hiprt.DeviceArray(np.zeros(1))
    for ctx, x, q, xd, qd in (big, small, big):
        want = ctx.beam_potential(x, q)
        e.write(np.array([-1.0]))
        _call(mmc, ctx, xd, qd, e)
        hiprt.check(hiprt.rt().hipDeviceSynchronize(), "hipDeviceSynchronize")
        got = float(e.read()[0])
        print("beams", ctx.n_beams, want, got)
        assert got == want


def test_noalloc_beam_potential_in_a_captured_graph(gpu, po):
    mmc = gpu
    ctx, x, q, xd, qd = _case(mmc, po, mmc.cubic_config(6, 5, 4), 6)
    want = ctx.beam_potential(x, q)
    e = hiprt.DeviceArray(np.zeros(1))
    s = hiprt.Stream()
    g = hiprt.Graph(s, lambda: _call(mmc, ctx, xd, qd, e, s.h))
    for r in range(3):
        e.write(np.array([-1.0]))
        g.replay()
        s.sync()
        assert float(e.read()[0]) == want, r
        # an eager call between the replays, on the same static scratch
        e.write(np.array([-1.0]))
        _call(mmc, ctx, xd, qd, e, s.h)
        s.sync()
        assert float(e.read()[0]) == want, r
    g.close()


def test_noalloc_rejects_a_foreign_batch(gpu):
    """A batched context is refused with the single-lattice API's error."""
    mmc = gpu
    sim = mmc.BatchSimulator([mmc.cubic_config(3, 3, 3), mmc.cubic_config(4, 3, 3)])
    e = hiprt.DeviceArray(np.zeros(1))
    x = hiprt.DeviceArray(np.zeros(3))
    rc = mmc.lib().mmc_beam_potential_noalloc(sim.context(), x.ptr, x.ptr, e.ptr, None)
    assert rc != 0
    assert "mmc_beam_potential_noalloc" in mmc.lib().mmc_last_error(sim.context()).decode()
