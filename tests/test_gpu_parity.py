"""GPU parity of the gfx950 hot path against the FP64 CPU oracle.

Every call goes through the C-ABI of include/mmc.h (libmmc.so).  The oracle is
the CPU restatement of MC/src/sim.cpp (pinned bit for bit against the reference
compiled from source in tests/test_oracle_pin.py).  Tolerances (FP64 path):
  objective energy and diagnostics      1e-12 relative
  gradient, trial records               1e-10 of the largest component
  (the device sums per node through the CSR incidence list and per block in a
  tree, the reference sums sequentially: only the summation order differs)
  converged equilibria (minimize, run)  the band the gradient tolerance
  leaves: two correct solvers stop anywhere with |g| <= 1e-6, and on the 8x8x4
  curve that band is ~1e-4 of the peak platen force
"""
import numpy as np
import pytest

from conftest import GOLDEN, rel, seeded_z

pytestmark = pytest.mark.gpu


def cases(mmc):
    return {
        "bcc_quick": mmc.quick_config(),
        "cubic_8x8x4": mmc.cubic_config(8, 8, 4),
        "cubic_4x5x3_jitter": mmc.cubic_config(4, 5, 3, sigma=1e-4, radius_mode=1, r_bot=3e-4, r_top=7e-4),
    }


def make_pair(mmc, po, cfg, plastic_seed=None, precision=0):
    o = po.Solver(cfg, "orc")
    lat = o.lattice()
    if plastic_seed is not None:
        rng = np.random.default_rng(plastic_seed)
        B = o.n_beams
        p = np.zeros((B, 7))
        p[:, 0] = rng.normal(0, 0.01, B)
        p[:, 1] = np.abs(rng.normal(0, 0.01, B))
        p[:, 2:5] = rng.normal(0, 0.02, (B, 3))
        p[:, 5] = np.abs(rng.normal(0, 0.01, B))
        p[:, 6] = np.abs(rng.normal(0, 0.01, B))
        o.set_plastic(p.reshape(-1))
        lat.plastic = p.reshape(-1).copy()
    ctx = mmc.Context(lat, cfg, precision)
    return o, ctx, lat


STATES = ["random", "squash", "yielded"]


@pytest.mark.parametrize("case", ["bcc_quick", "cubic_8x8x4", "cubic_4x5x3_jitter"])
@pytest.mark.parametrize("state", STATES)
def test_objective_parity(gpu, po, case, state):
    mmc = gpu
    cfg = cases(mmc)[case]
    o, ctx, lat = make_pair(mmc, po, cfg, plastic_seed=7 if state == "yielded" else None)
    y = lat.x.reshape(-1, 3)[:, 1]
    jr = cfg.joint_scale * cfg.max_radius
    z = seeded_z(o.n_nodes, 11, squash=0.8 if state == "squash" else 0.0, y=y, jr=jr, ts=cfg.cell_size)
    top = o.top0 - 0.02 * o.height
    ro, go, to = o.objective(z, top, trials=True)
    rg, gg = ctx.objective(z, top, trials=True)
    tg = ctx.read_trial()
    print(case, state, "contacts", ro.contacts, rg.contacts, "E", ro.energy, rg.energy)
    assert rg.finite == 1
    assert rg.contacts == ro.contacts
    assert abs(rg.energy - ro.energy) <= 1e-12 * abs(ro.energy)
    for k in ("elastic", "plastic", "contact", "reaction"):
        a, b = getattr(rg, k), getattr(ro, k)
        assert abs(a - b) <= 1e-12 * max(abs(b), 1e-300) + 1e-300, (k, a, b)
    assert abs(rg.max_pen - ro.max_pen) <= 1e-15
    assert rel(gg, go) <= 1e-10
    assert abs(rg.rms - ro.rms) <= 1e-10 * ro.rms
    assert abs(rg.max - ro.max) <= 1e-10 * ro.max
    for c in range(14):
        assert rel(tg[:, c], to[:, c], 1e-300) <= 1e-10, c
    assert np.array_equal(tg[:, 14:], to[:, 14:])


def test_objective_state_kinds(gpu, po):
    """The squashed states exercise sphere-sphere, sphere-capsule and
    capsule-capsule contact and the yielded state all three return maps."""
    mmc = gpu
    cfg = cases(mmc)["cubic_8x8x4"]
    o, ctx, lat = make_pair(mmc, po, cfg, plastic_seed=3)
    y = lat.x.reshape(-1, 3)[:, 1]
    z = seeded_z(o.n_nodes, 5, r_sigma=0.4, squash=0.88, y=y, jr=cfg.joint_scale * cfg.max_radius,
                 ts=cfg.cell_size)
    rg, _ = ctx.objective(z, o.top0 - 0.1 * o.height, trials=True)
    ro, _ = o.objective(z, o.top0 - 0.1 * o.height)
    t = ctx.read_trial()
    assert rg.contacts == ro.contacts and rg.contacts > 100
    assert t[:, 14].sum() > 0 and t[:, 15].sum() > 0 and t[:, 16].sum() > 0


def test_objective_locked_and_inf(gpu, po):
    """Locked DOFs carry zero gradient; a collapsed configuration with a
    non-finite energy reports INF and zero norms (MC/src/sim.cpp:172)."""
    mmc = gpu
    cfg = cases(mmc)["cubic_8x8x4"]
    o, ctx, lat = make_pair(mmc, po, cfg)
    z = seeded_z(o.n_nodes, 2)
    _, g = ctx.objective(z, o.top0)
    assert np.all(g[lat.locked.astype(bool)] == 0)
    zbad = z.copy()
    zbad[6 * 10 + 1] = np.inf
    rg, gb = ctx.objective(zbad, o.top0)
    assert rg.finite == 0 and np.isinf(rg.energy) and np.all(gb == 0)


def test_minimize_commit_parity(gpu, po, loop_engine):
    """Both solvers stop anywhere inside the gradient tolerance (1e-6): the
    equilibria agree to that band, not to round-off (force to 1e-5; positions
    to 5e-6 relative, measured 1.5e-6 at most over two summation orders of the
    GPU's reductions)."""
    mmc = gpu
    cfg = cases(mmc)["cubic_8x8x4"]
    o, ctx, lat = make_pair(mmc, po, cfg)
    for s in (0.0, 0.01, 0.02):
        top = o.top0 - s * o.height
        so, fo, zo = o.minimize(top)
        sg, fg, zg = ctx.minimize(top)
        print(s, so.iterations, sg.iterations, fo.reaction, fg.reaction)
        assert so.converged and sg.converged
        assert sg.rms <= cfg.gradient_rms_tolerance and sg.max <= cfg.gradient_max_tolerance
        assert abs(fg.reaction - fo.reaction) <= 1e-5 * max(fo.reaction, 1.0)
        assert abs(fg.energy - fo.energy) <= 1e-8 * abs(fo.energy)
        o.commit()
        ctx.commit()
        xo, qo, po_ = o.read_state()
        xg, qg, pg = ctx.read_state()
        assert rel(xg, xo) <= 5e-6
        assert np.max(np.abs(pg - po_)) <= 2e-4 * max(1e-3, np.max(np.abs(po_)))


def test_minimize_deterministic(gpu, po, loop_engine):
    mmc = gpu
    cfg = cases(mmc)["bcc_quick"].replace(relative_escape=0, gradient_rms_tolerance=1e-5,
                                           gradient_max_tolerance=1e-4, max_iterations=400)
    outs = []
    for _ in range(2):
        o, ctx, lat = make_pair(mmc, po, cfg)
        st, fin, z = ctx.minimize(o.top0 - 0.01 * o.height)
        outs.append((st.iterations, fin.energy, z.copy()))
    assert outs[0][0] == outs[1][0]
    assert outs[0][1] == outs[1][1]
    assert np.array_equal(outs[0][2], outs[1][2])


def test_commit_requires_trial(gpu, po):
    mmc = gpu
    o, ctx, lat = make_pair(mmc, po, cases(mmc)["cubic_4x5x3_jitter"])
    with pytest.raises(mmc.MMCError, match="missing constitutive trial states"):
        ctx.commit()


def test_state_stats(gpu, po):
    mmc = gpu
    o, ctx, lat = make_pair(mmc, po, cases(mmc)["cubic_8x8x4"], plastic_seed=9)
    s = ctx.stats()
    p = lat.plastic.reshape(-1, 7)
    pa, pr = np.abs(p[:, 0]), np.linalg.norm(p[:, 2:5], axis=1)
    assert abs(s.sum_axial - pa.sum()) <= 1e-12 * pa.sum()
    assert s.max_axial == pa.max() and abs(s.max_rot - pr.max()) <= 1e-15
    assert s.plastic_beams == int((pa > 1e-5).sum()) and s.plastic_hinges == int((pr > 1e-4).sum())
    y = lat.x.reshape(-1, 3)[:, 1]
    assert s.y_min == y.min() and s.y_max == y.max()


def test_beam_potential_and_checks(gpu, po):
    mmc = gpu
    cfg = mmc.quick_config()
    o = po.Solver(cfg, "orc")
    sim = mmc.Simulator(cfg)
    ok, err = sim.rigid_rotation_test()
    oko, erro = o.rigid_rotation_test()
    assert ok and oko and err < 1e-12
    ok, err = sim.gradient_check()
    oko, erro = o.gradient_check()
    print("fd", err, erro)
    assert ok and abs(err - erro) <= 0.05 * erro
    lat = o.lattice()
    ctx = mmc.Context(lat, cfg)
    rng = np.random.default_rng(4)
    x = lat.x + rng.normal(0, 1e-4, lat.x.shape)
    q = lat.q.reshape(-1, 4) + rng.normal(0, 0.01, (o.n_nodes, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).reshape(-1)
    assert abs(ctx.beam_potential(x, q) - o.beam_potential(x, q)) <= 1e-12 * o.beam_potential(x, q)


def _frames_error(fg, fo):
    peak = max(float(np.max(fo["reaction_force"])), 1e-12)
    return {
        "force": float(np.max(np.abs(fg["reaction_force"] - fo["reaction_force"])) / peak),
        "loading_work": abs(fg["loading_work"][-1] - fo["loading_work"][-1]) / max(fo["loading_work"][-1], 1e-12),
        "dissipated": abs(fg["dissipated_work"][-1] - fo["dissipated_work"][-1]) / max(fo["dissipated_work"][-1], 1e-12),
        "permanent_set": abs(fg["permanent_set"][-1] - fo["permanent_set"][-1]),
    }


def test_run_config1_parity(gpu, po, loop_engine):
    """Full 60 + 60 step schedule of the 8x8x4 cubic lattice (plain FP64 GPU
    path) against the oracle's force-displacement curve."""
    mmc = gpu
    cfg = mmc.cubic_config(8, 8, 4)
    path = GOLDEN / "cubic_8x8x4_oracle_frames.npy"
    fo = np.load(path) if path.exists() else po.Solver(cfg, "orc").run()
    sim = mmc.Simulator(cfg)
    fg = sim.run()
    assert len(fg) == len(fo)
    err = _frames_error(fg, fo)
    print("config1 errors", err, "iterations", fg["iterations"].sum(), fo["iterations"].sum())
    assert np.all(fg["converged"] == 1)
    peak = fo["reaction_force"].max()
    per_frame = np.abs(fg["reaction_force"] - fo["reaction_force"]) / peak
    print("worst frames", np.argsort(per_frame)[-5:], np.sort(per_frame)[-5:])
    # gradient tolerance 1e-6 bounds the agreement of two correct FP64 solvers
    assert err["force"] <= 2e-4
    assert err["loading_work"] <= 2e-5 and err["dissipated"] <= 5e-5


def test_device_loop_matches_host_loop(gpu, po, loop_engine, monkeypatch):
    """The graph-replayed phase machine and the host-driven loop implement the
    same algorithm: same convergence, equilibria inside the tolerance band.
    The two sum in different orders, so their iterates part within the band
    the gradient tolerance (1e-6) leaves; each commit carries that difference
    into the next increment's plastic state, so the stored energy after four
    increments agrees to ~1e-6 relative (measured 2e-7 to 1.1e-6), the platen
    force to ~2e-6."""
    mmc = gpu
    cfg = cases(mmc)["cubic_8x8x4"]
    res = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("MMC_HOST_LOOP", mode)
        o, ctx, lat = make_pair(mmc, po, cfg)
        out = []
        for s in (0.0, 0.01, 0.02, 0.03):
            st, fin, z = ctx.minimize(o.top0 - s * o.height)
            ctx.commit()
            out.append((st.converged, st.iterations, fin.reaction, fin.energy))
        res[mode] = out
    print(res)
    for a, b in zip(res["0"], res["1"]):
        assert a[0] == b[0] == 1
        assert abs(a[2] - b[2]) <= 1e-5 * max(abs(b[2]), 1.0)
        assert abs(a[3] - b[3]) <= 5e-6 * abs(b[3])
        assert abs(a[1] - b[1]) <= 0.3 * b[1] + 5


# ---------------------------------------------------------------------------
# Per-term energy line search (line_search = 1) and the mixed-precision policy
# (FP32 shear/rotation element arithmetic, FP64 axial strain, positions and
# reductions, approximate-Wolfe acceptance below the FP32 energy noise)
# ---------------------------------------------------------------------------
def test_per_term_line_search_matches_oracle(gpu, po, loop_engine):
    """line_search = 1 on both sides, at a gradient tolerance (1e-7) below the
    total-energy Armijo floor of the 8x8x4 lattice: both converge, and agree
    to the tighter band."""
    mmc = gpu
    cfg = cases(mmc)["cubic_8x8x4"].replace(line_search=1, gradient_rms_tolerance=1e-7,
                                            gradient_max_tolerance=1e-7, relative_escape=0)
    o, ctx, lat = make_pair(mmc, po, cfg)
    for s in (0.0, 0.01, 0.02):
        top = o.top0 - s * o.height
        so, fo, zo = o.minimize(top)
        sg, fg, zg = ctx.minimize(top)
        print(s, so.iterations, sg.iterations, fo.reaction, fg.reaction)
        assert so.converged and sg.converged
        assert sg.rms <= 1e-7 and sg.max <= 1e-7
        assert abs(fg.reaction - fo.reaction) <= 2e-6 * max(fo.reaction, 1.0)
        o.commit()
        ctx.commit()


@pytest.mark.parametrize("case", ["cubic_8x8x4", "cubic_4x5x3_jitter"])
@pytest.mark.parametrize("state", STATES)
def test_objective_parity_mixed(gpu, po, case, state):
    """Mixed-precision objective against the FP64 oracle: rotations and shear
    in FP32 bound the agreement (energy 1e-7 relative, gradient 2e-5 of its
    largest component); contacts and axial terms stay FP64."""
    mmc = gpu
    cfg = cases(mmc)[case]
    o, ctx, lat = make_pair(mmc, po, cfg, plastic_seed=7 if state == "yielded" else None,
                            precision=mmc.MMC_MIXED)
    y = lat.x.reshape(-1, 3)[:, 1]
    jr = cfg.joint_scale * cfg.max_radius
    z = seeded_z(o.n_nodes, 11, squash=0.8 if state == "squash" else 0.0, y=y, jr=jr, ts=cfg.cell_size)
    top = o.top0 - 0.02 * o.height
    ro, go, to = o.objective(z, top, trials=True)
    rg, gg = ctx.objective(z, top, trials=True)
    e_rel = abs(rg.energy - ro.energy) / abs(ro.energy)
    g_rel = rel(gg, go)
    print(case, state, "energy rel", e_rel, "grad rel", g_rel)
    assert rg.finite == 1 and rg.contacts == ro.contacts
    assert e_rel <= 1e-7
    assert g_rel <= 2e-5
    assert abs(rg.reaction - ro.reaction) <= 1e-12 * max(ro.reaction, 1.0)


def test_mixed_minimize_matches_fp64(gpu, po, loop_engine):
    """Mixed-precision equilibria against the FP64 oracle at gradient
    tolerance 1e-7: platen force within 1e-5 relative."""
    mmc = gpu
    cfg = cases(mmc)["cubic_8x8x4"].replace(line_search=1, gradient_rms_tolerance=1e-7,
                                            gradient_max_tolerance=1e-7, relative_escape=0)
    o, ctx, lat = make_pair(mmc, po, cfg, precision=mmc.MMC_MIXED)
    for s in (0.0, 0.01, 0.02, 0.03):
        top = o.top0 - s * o.height
        so, fo, zo = o.minimize(top)
        sg, fg, zg = ctx.minimize(top)
        print(s, so.iterations, sg.iterations, fo.reaction, fg.reaction, sg.rms)
        assert so.converged and sg.converged
        assert abs(fg.reaction - fo.reaction) <= 1e-5 * max(fo.reaction, 1.0)
        o.commit()
        ctx.commit()


def test_run_config1_mixed(gpu, po, loop_engine):
    """Full 60 + 60 schedule of the 8x8x4 lattice with the mixed policy against
    the oracle's FP64 curve (same band as the FP64 path at tolerance 1e-6)."""
    mmc = gpu
    cfg = mmc.cubic_config(8, 8, 4)
    path = GOLDEN / "cubic_8x8x4_oracle_frames.npy"
    fo = np.load(path) if path.exists() else po.Solver(cfg, "orc").run()
    fg = mmc.Simulator(cfg, mmc.MMC_MIXED).run()
    err = _frames_error(fg, fo)
    print("config1 mixed errors", err, "iterations", fg["iterations"].sum(), fo["iterations"].sum())
    assert np.all(fg["converged"] == 1)
    assert err["force"] <= 2e-4
    assert err["loading_work"] <= 2e-5 and err["dissipated"] <= 5e-5


@pytest.mark.parametrize("split", ["0", "1", "2", "4", "16"])
def test_launch_shapes_leave_results_unchanged(gpu, split, monkeypatch):
    """The loop's tuner picks the decision kernel's split per context, or
    fuses the decision into the gather kernel (split 0); it may not change a
    result: the 8x8x4 curve (mixed policy, 12 + 4 increments) with each fixed
    shape equals, bit for bit, the curve at the untuned split (8 blocks)."""
    mmc = gpu
    cfg = mmc.cubic_config(8, 8, 4, load=12, unload=4)
    monkeypatch.setenv("MMC_DEVICE_LOOP", "graph")
    monkeypatch.setenv("MMC_TUNE", "0")
    ref = mmc.Simulator(cfg, mmc.MMC_MIXED).run()
    monkeypatch.delenv("MMC_TUNE")
    monkeypatch.setenv("MMC_DECIDE_BLOCKS", split)
    sim = mmc.Simulator(cfg, mmc.MMC_MIXED)
    got = sim.run()
    d = mmc.C.c_int32()
    mmc.lib().mmc_loop_shape(sim.context(), mmc.C.byref(d), None)
    assert d.value == int(split)
    assert len(got) == len(ref)
    for name in ref.dtype.names:
        assert np.array_equal(got[name], ref[name]), name


@pytest.mark.parametrize("serial", ["-1", "0", "1000000"])
def test_contact_routing_leaves_results_unchanged(gpu, serial, monkeypatch):
    """The graph loop evaluates a chunk of 64 contact shapes either one lane
    per shape or, when its longest Verlet list is long, four lanes per shape
    whose pair contributions are added in ascending slot order.  Routing every
    chunk to the lane groups (-1), every chunk with a partner (0) or none
    (1000000) must give the default curve bit for bit: 8x8x4 at 30% strain
    (platen and self-contact), mixed policy, 12 + 4 increments."""
    mmc = gpu
    cfg = mmc.cubic_config(8, 8, 4, load=12, unload=4)
    monkeypatch.setenv("MMC_DEVICE_LOOP", "graph")
    ref = mmc.Simulator(cfg, mmc.MMC_MIXED).run()
    assert ref["active_contacts"].max() > 0
    monkeypatch.setenv("MMC_CONTACT_SERIAL", serial)
    got = mmc.Simulator(cfg, mmc.MMC_MIXED).run()
    assert len(got) == len(ref)
    for name in ref.dtype.names:
        assert np.array_equal(got[name], ref[name]), name
