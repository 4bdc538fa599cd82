"""Closed-form and brute-force checks of the oracle's building blocks.

These pin what the fixtures alone would leave open, with checks chosen so that a
dropped term, a wrong sign or a transposed operand fails:
  * return maps: elastic/plastic branches, consistency, variational property
    (d/dx of the incremental potential equals the returned force)
  * SO(3): exp/log inverse pair, Rodrigues rotation, J_l^{-T} against the
    inverse transpose of the closed-form left Jacobian
  * segment closest points: brute-force (s, t) sampling
  * objective: exact translational gradient by finite differences, rigid-body
    invariances giving gravity and centering energies, platen reaction by
    direct summation, and contact energy by an O(n^2) all-pairs narrowphase
"""
import numpy as np
import pytest

from conftest import seeded_z


def skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


def rodrigues(v):
    a = np.linalg.norm(v)
    K = skew(v / a)
    return np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K


def quat_to_mat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


@pytest.fixture(params=["orc", "ref"])
def prefix(request, po):
    if not po.available(request.param):
        pytest.skip(f"{request.param} library not built")
    return request.param


def R1(po, prefix, *a):
    o = po.helper(prefix, "scalar_return", *a)
    return dict(force=o[0], p=o[1], a=o[2], dg=o[3], energy=o[4], hard=o[5], yielded=o[6])


def test_scalar_return_branches(po, prefix):
    k, y, H = 2e9, 5e7, 1e8
    r = R1(po, prefix, 0.01, 0.0, 0.0, k, y, H)  # trial stress 2e7 < yield: elastic
    assert r["yielded"] == 0 and r["p"] == 0 and r["a"] == 0 and r["dg"] == 0
    assert r["force"] == pytest.approx(k * 0.01, rel=1e-15)
    assert r["energy"] == pytest.approx(0.5 * k * 0.01 ** 2, rel=1e-15)
    for x, pc, ac in [(0.05, 0.0, 0.0), (-0.07, 0.01, 0.003), (0.2, -0.02, 0.05)]:
        r = R1(po, prefix, x, pc, ac, k, y, H)
        f = abs(k * (x - pc)) - (y + H * ac)
        assert r["yielded"] == 1
        assert r["dg"] == pytest.approx(f / (k + H), rel=1e-14)
        # consistency: the returned stress sits on the hardened yield surface
        assert abs(r["force"]) == pytest.approx(y + H * r["a"], rel=1e-12)
        assert np.sign(r["force"]) == np.sign(k * (x - pc))
        assert r["hard"] == pytest.approx(0.5 * H * r["a"] ** 2, rel=1e-15)


@pytest.mark.parametrize("x,pc,ac", [(0.05, 0.0, 0.0), (-0.07, 0.01, 0.003), (0.004, 0.0, 0.0)])
def test_scalar_return_variational(po, prefix, x, pc, ac):
    """W(x) = energy + hard + y*dg is the incremental potential: dW/dx = force."""
    k, y, H = 2e9, 5e7, 1e8

    def W(v):
        r = R1(po, prefix, v, pc, ac, k, y, H)
        return r["energy"] + r["hard"] + y * r["dg"]
    h = 1e-7
    num = (W(x + h) - W(x - h)) / (2 * h)
    assert num == pytest.approx(R1(po, prefix, x, pc, ac, k, y, H)["force"], rel=1e-6)


def test_vector_return(po, prefix):
    k, y, H = 0.02, 1e-3, 0.0012
    rng = np.random.default_rng(0)
    for _ in range(20):
        t = rng.normal(0, 0.2, 3)
        pc = rng.normal(0, 0.02, 3)
        ac = abs(rng.normal(0, 0.01))
        o = po.helper(prefix, "vector_return", t, pc, ac, k, y, H)
        F, p, a, dg = o[0:3], o[3:6], o[6], o[7]
        trial = k * (t - pc)
        f = np.linalg.norm(trial) - (y + H * ac)
        if f > 0:
            assert o[10] == 1 and dg == pytest.approx(f / (k + H), rel=1e-13)
            assert np.linalg.norm(F) == pytest.approx(y + H * a, rel=1e-12)
            # radial return: plastic flow along the trial direction
            assert np.allclose(np.cross(p - pc, trial), 0, atol=1e-15)
        else:
            assert o[10] == 0 and np.array_equal(p, pc)
        W = lambda v: (lambda r: r[8] + r[9] + y * r[7])(  # noqa: E731
            po.helper(prefix, "vector_return", v, pc, ac, k, y, H))
        h = 1e-7
        num = np.array([(W(t + h * e) - W(t - h * e)) / (2 * h) for e in np.eye(3)])
        assert np.allclose(num, F, rtol=1e-5, atol=1e-12)


def test_so3_helpers(po, prefix):
    rng = np.random.default_rng(1)
    for _ in range(30):
        v = rng.normal(0, 1.0, 3)
        v *= min(1.0, 3.0 / np.linalg.norm(v))
        q = po.helper(prefix, "quat_exp", v)
        assert np.linalg.norm(q) == pytest.approx(1.0, abs=1e-15)
        assert np.allclose(quat_to_mat(q), rodrigues(v), atol=1e-14)
        assert np.allclose(po.helper(prefix, "quat_log", q), v, atol=1e-13)
        # J_l^{-T}: inverse transpose of J_l = I + (1-cos a)/a^2 K + (a - sin a)/a^3 K^2
        a = np.linalg.norm(v)
        K = skew(v)
        Jl = np.eye(3) + (1 - np.cos(a)) / a ** 2 * K + (a - np.sin(a)) / a ** 3 * K @ K
        w = rng.normal(0, 1, 3)
        assert np.allclose(po.helper(prefix, "jl_inv_t", v, w), np.linalg.inv(Jl).T @ w, rtol=1e-12, atol=1e-13)
    small = np.array([1e-9, -2e-9, 3e-9])
    w = np.array([0.3, -0.2, 0.5])
    assert np.allclose(po.helper(prefix, "jl_inv_t", small, w), w + 0.5 * np.cross(small, w), atol=1e-15)
    assert np.allclose(po.helper(prefix, "quat_exp", small), [1, *(0.5 * small)], atol=1e-20)


def test_closest_segments_bruteforce(po, prefix):
    rng = np.random.default_rng(2)
    S = np.linspace(0, 1, 801)
    cases = [rng.normal(0, 1, (4, 3)) for _ in range(12)]
    base = rng.normal(0, 1, 3)
    d = rng.normal(0, 1, 3)
    cases.append(np.array([base, base + d, base + [0, 0.5, 0], base + [0, 0.5, 0] + 2 * d]))  # parallel
    cases.append(np.array([base, base, base + [1, 0, 0], base + [1, 1, 0]]))  # point vs segment
    for p1, q1, p2, q2 in cases:
        o = po.helper(prefix, "closest_segments", p1, q1, p2, q2)
        s, t = o[0], o[1]
        assert 0 <= s <= 1 and 0 <= t <= 1
        assert np.allclose(o[2:5], p1 + s * (q1 - p1), atol=1e-14)
        assert np.allclose(o[5:8], p2 + t * (q2 - p2), atol=1e-14)
        dist = np.linalg.norm(o[5:8] - o[2:5])
        A = p1 + S[:, None] * (q1 - p1)
        B = p2 + S[:, None] * (q2 - p2)
        brute = np.sqrt(((A[:, None, :] - B[None, :, :]) ** 2).sum(-1)).min()
        assert dist <= brute + 1e-12
        assert dist >= brute - 2e-3 * (np.linalg.norm(q1 - p1) + np.linalg.norm(q2 - p2))


def small_cubic(mmc):
    return mmc.cubic_config(3, 3, 2)


def test_translational_gradient_exact(mmc, po):
    """The translational gradient is the exact derivative of the energy (the
    rotational one is the reference's first-order corotational tangent)."""
    cfg = small_cubic(mmc)
    s = po.Solver(cfg, "orc")
    z = seeded_z(s.n_nodes, 4, t_sigma=0.002, r_sigma=0.002)
    top = s.top0 + 0.01
    out, g = s.objective(z, top)
    rng = np.random.default_rng(5)
    lk = s.lattice().locked.astype(bool)
    dofs = [k for k in rng.choice(6 * s.n_nodes, 60, replace=False) if k % 6 < 3 and not lk[k]][:12]
    for k in dofs:
        h = 1e-7
        zp, zm = z.copy(), z.copy()
        zp[k] += h
        zm[k] -= h
        num = (s.objective(zp, top)[0].energy - s.objective(zm, top)[0].energy) / (2 * h)
        assert num == pytest.approx(g[k], rel=2e-5, abs=1e-9 * np.abs(g).max())


def test_rotational_gradient_tangent(mmc, po):
    s = po.Solver(small_cubic(mmc), "orc")
    ok, err = s.gradient_check()
    assert ok and err < 0.075


def test_gravity_and_centering_energies(mmc, po):
    """Rigid translations leave beam energy unchanged: a vertical shift changes
    the energy by sum(m g dy), a horizontal one by the centering term."""
    cfg = small_cubic(mmc)
    s = po.Solver(cfg, "orc")
    L = s.lattice()
    lk = L.locked.reshape(-1, 6).astype(bool)
    N, ts = s.n_nodes, cfg.cell_size
    top = s.top0 + 1.0
    z0 = np.zeros((N, 6))
    z0[:, 1] = 0.2  # lift off the lower platen
    e0 = s.objective(z0.reshape(-1), top)[0].energy
    dy = 0.01
    z1 = z0.copy()
    z1[:, 1] += dy
    e1 = s.objective(z1.reshape(-1), top)[0].energy
    assert e1 - e0 == pytest.approx(float(L.mass.sum()) * cfg.gravity * dy * ts, rel=1e-9)
    # horizontal shift: locked x/z DOFs stay, so shift only the free nodes and
    # compare with the centering energy of the resulting centroid
    dx = 0.01
    z2 = z0.copy()
    z2[:, 0] += np.where(lk[:, 0], 0.0, dx)
    z2[:, 2] += np.where(lk[:, 2], 0.0, dx)
    x0 = L.x.reshape(-1, 3)
    x2 = x0 + ts * np.where(lk[:, :3], 0, z2[:, :3])
    c0, c2 = x0.mean(0), x2.mean(0)
    cen = lambda c: 0.5 * cfg.centering_stiffness * (c[0] ** 2 + c[2] ** 2)  # noqa: E731
    # beams between a locked and a free node stretch: subtract their energy change
    e2 = s.objective(z2.reshape(-1), top)[0]
    e2b = s.objective(z0.reshape(-1), top)[0]
    beam_change = e2.elastic - e2b.elastic
    assert (e2.energy - e0) - beam_change == pytest.approx(cen(c2) - cen(c0), rel=1e-6, abs=1e-15)


def test_platen_reaction_direct(mmc, po):
    cfg = small_cubic(mmc)
    s = po.Solver(cfg, "orc")
    L = s.lattice()
    z = seeded_z(s.n_nodes, 8, t_sigma=0.01, r_sigma=0.01)
    lk = L.locked.reshape(-1, 6).astype(bool)
    x = L.x.reshape(-1, 3) + cfg.cell_size * np.where(lk[:, :3], 0, z.reshape(-1, 6)[:, :3])
    jr = cfg.joint_scale * cfg.max_radius
    for top in (s.top0 - 0.003, s.top0 - 0.006):
        out, _ = s.objective(z, top)
        pt = np.maximum(0, x[:, 1] + jr - top)
        assert out.reaction == pytest.approx(cfg.contact_stiffness * pt.sum(), rel=1e-13)


def test_contact_all_pairs(mmc, po):
    """Hashed broadphase + exclusion == O(n^2) all-pairs narrowphase."""
    cfg = mmc.cubic_config(3, 3, 2)
    s = po.Solver(cfg, "orc")
    L = s.lattice()
    N, B = s.n_nodes, s.n_beams
    z = seeded_z(N, 9, t_sigma=0.05, squash=0.88, y=L.x.reshape(-1, 3)[:, 1],
                 jr=cfg.joint_scale * cfg.max_radius, ts=cfg.cell_size)
    top = s.top0 + 1.0  # no upper platen contact
    out, _ = s.objective(z, top)
    lk = L.locked.reshape(-1, 6).astype(bool)
    x = L.x.reshape(-1, 3) + cfg.cell_size * np.where(lk[:, :3], 0, z.reshape(-1, 6)[:, :3])
    jr = cfg.joint_scale * cfg.max_radius
    conn = set(int(c) for c in L.connected)
    key = lambda u, v: (min(u, v) << 32) | max(u, v)  # noqa: E731
    shapes = [(x[i], x[i], jr, (i,)) for i in range(N)] + \
             [(x[L.a[k]], x[L.b[k]], L.radius[k], (int(L.a[k]), int(L.b[k]))) for k in range(B)]
    energy, count = 0.0, 0
    kc, margin = cfg.contact_stiffness, cfg.contact_margin
    for i in range(len(shapes)):
        for j in range(i + 1, len(shapes)):
            A, Bs = shapes[i], shapes[j]
            if any(u == v or key(u, v) in conn for u in A[3] for v in Bs[3]):
                continue
            o = po.helper("orc", "closest_segments", A[0], A[1], Bs[0], Bs[1])
            dist = np.linalg.norm(o[5:8] - o[2:5])
            target = A[2] + Bs[2] + margin
            if dist < target:
                energy += 0.5 * kc * (target - dist) ** 2
                count += 1
    bottom = np.maximum(0, jr - x[:, 1])
    energy += 0.5 * kc * (bottom ** 2).sum()
    count += int((bottom > 0).sum())
    assert count > 10 and out.contacts == count
    assert out.contact == pytest.approx(energy, rel=1e-11)


def test_per_term_energy_change_is_the_energy_difference(po):
    """line_search 1 sums the energy change per beam, node and contact shape.
    Pinned two ways: for a finite step it equals the difference of the total
    energies to round-off; for a tiny translational step (where the gradient
    is the exact derivative) it follows the first-order change g.dz far below
    the round-off floor of the totals."""
    import mmc
    from conftest import seeded_z
    cfg = mmc.cubic_config(4, 4, 3)
    o = po.Solver(cfg, "orc")
    lat = o.lattice()
    y = lat.x.reshape(-1, 3)[:, 1]
    jr = cfg.joint_scale * cfg.max_radius
    top = o.top0 - 0.03 * o.height
    z0 = seeded_z(o.n_nodes, 5, squash=0.6, y=y, jr=jr, ts=cfg.cell_size)
    dz = np.random.default_rng(9).normal(0, 1e-3, (o.n_nodes, 6))
    e0, g0 = o.objective(z0, top)[:2]
    per, tot = o.energy_change(z0, z0 + dz.reshape(-1), top)
    assert abs(per - tot) <= 1e-12 * abs(e0.energy)
    dz[:, 3:] = 0.0  # translations: the gradient there is exact
    dz = dz.reshape(-1)
    h = 1e-9
    per_h, tot_h = o.energy_change(z0, z0 + h * dz, top)
    lin = h * float(g0 @ dz)
    assert abs(per_h - lin) <= 2e-4 * abs(lin)  # measured 3.4e-5
    assert abs(per_h - lin) * 5 < abs(tot_h - lin)  # total difference: 7e-4


def test_approximate_wolfe_line_search_reaches_the_same_equilibrium(mmc, po):
    """line_search 2 completes the per-term Armijo test with the approximate-
    Wolfe slope test of Hager & Zhang below 1e-6 |E|.  Pinned against the
    per-term test alone on increments both converge (the equilibrium is the
    same: platen force to 1e-7 relative, energy to 1e-10), and by what it adds:
    at a gradient tolerance of 3e-9, below the per-term test's resolution on
    this lattice, line_search 2 still converges while line_search 1 stops at
    a line-search failure."""
    base = mmc.cubic_config(4, 4, 3)
    top = None
    res = {}
    for ls in (1, 2):
        cfg = base.replace(line_search=ls, gradient_rms_tolerance=1e-7, gradient_max_tolerance=1e-7,
                           relative_escape=0)
        o = po.Solver(cfg, "orc")
        top = o.top0 - 0.05 * o.height
        st, fin, _ = o.minimize(top)
        assert st.converged
        res[ls] = fin
    assert abs(res[2].reaction - res[1].reaction) <= 1e-7 * res[1].reaction
    assert abs(res[2].energy - res[1].energy) <= 1e-10 * abs(res[1].energy)
    deep = {}
    for ls in (1, 2):
        cfg = base.replace(line_search=ls, gradient_rms_tolerance=3e-9, gradient_max_tolerance=3e-9,
                           relative_escape=0, max_iterations=20000)
        o = po.Solver(cfg, "orc")
        st, fin, _ = o.minimize(top)
        deep[ls] = st
    assert deep[2].converged and deep[2].max <= 3e-9
    assert not deep[1].converged


def test_capture_closed_forms(mmc, po):
    """Simulator::capture (MC/src/sim.cpp:215-310) against what its bodies
    must be: a strut's orientation turns the x axis onto its chord (alignX)
    with a proper rotation, it sits at the chord's midpoint with half extents
    (L/2, r, r); a joint's velocity is its displacement over the step time and
    its angular velocity the rotation vector it turned through over the step
    time; a strut moves with the mean of its two joints; the material class
    follows the plastic severity thresholds; the machine closes the frame with
    the moving platen at the frame's top."""
    cfg = mmc.default_config().replace(cells_x=2, cells_y=3, cells_z=2)
    s = po.Solver(cfg, "orc")
    lat = s.lattice()
    N, B = s.n_nodes, s.n_beams
    rng = np.random.default_rng(5)
    x0 = lat.x.reshape(N, 3)
    dx = rng.normal(0, 3e-4, (N, 3))
    rho = rng.normal(0, 0.2, (N, 3))
    q0 = lat.q.reshape(N, 4)
    assert np.allclose(q0, [1, 0, 0, 0])  # the lattice starts unrotated
    q1 = np.array([po.helper("orc", "quat_exp", r) for r in rho])
    sev = np.array([0.0, 0.2, 0.3, 0.5, 0.7, 1.0, 1.2])[np.arange(B) % 7]
    plastic = np.zeros((B, 7))
    plastic[:, 0] = 0.10 * sev * np.where(np.arange(B) % 2, 1, -1)
    s.set_pose((x0 + dx).reshape(-1), q1.reshape(-1))
    s.set_plastic(plastic.reshape(-1))
    top = s.top0 - 0.004
    b = s.capture(top, lat.x, lat.q)
    dt = cfg.step_time
    x1 = x0 + dx
    assert len(b) == B + N + 7
    # struts
    a_, b_ = lat.a, lat.b
    d = x1[b_] - x1[a_]
    L = np.linalg.norm(d, axis=1)
    for k in range(B):
        Rm = quat_to_mat(b["q"][k])
        assert np.allclose(Rm @ [1, 0, 0], d[k] / L[k], atol=1e-14)
        assert np.allclose(Rm.T @ Rm, np.eye(3), atol=1e-14) and np.linalg.det(Rm) > 0
    assert np.allclose(b["p"][:B], 0.5 * (x1[a_] + x1[b_]), rtol=0, atol=1e-17)
    assert np.allclose(b["half"][:B], np.c_[0.5 * L, lat.radius, lat.radius], rtol=1e-15, atol=0)
    assert np.array_equal(b["material"][:B], np.searchsorted([0.28, 0.62, 1.05], sev, side="right"))
    v = dx / dt
    assert np.allclose(b["v"][:B], 0.5 * (v[a_] + v[b_]), rtol=1e-12, atol=1e-15)
    assert np.allclose(b["w"][:B], 0.5 * (rho[a_] + rho[b_]) / dt, rtol=1e-10, atol=1e-12)
    assert np.array_equal(b["id"][:B], np.arange(B)) and np.array_equal(b["slot"][:B], np.arange(B))
    assert np.array_equal(b["kind"][:B], np.zeros(B)) and np.array_equal(b["ring"][:B], lat.beam_layer)
    assert np.array_equal(b["module_class"][:B], lat.beam_family)
    # joints
    j = b[B:B + N]
    jr = cfg.joint_scale * cfg.max_radius
    assert np.array_equal(j["p"], x1) and np.array_equal(j["q"], q1)
    assert np.allclose(j["v"], v, rtol=1e-12, atol=1e-15)
    assert np.allclose(j["w"], rho / dt, rtol=1e-10, atol=1e-12)
    assert np.all(j["half"] == jr) and np.all(j["material"] == 4) and np.all(j["kind"] == 1)
    assert np.array_equal(j["id"], B + np.arange(N)) and np.array_equal(j["ring"], lat.node_layer)
    # machine
    m = b[B + N:]
    assert np.array_equal(m["kind"], [2, 2, 3, 3, 3, 3, 3]) and np.array_equal(m["id"], B + N + np.arange(7))
    assert m["p"][1][1] == top + 0.0030 and m["p"][6][1] == top + 0.0080 and m["p"][0][1] == -0.0030
    assert np.all(m["q"] == [1, 0, 0, 0]) and np.all(m["v"] == 0) and np.all(m["slot"] == -1)
    # first frame: no previous pose, no motion
    f = s.capture(top)
    assert np.all(f["v"] == 0) and np.all(f["w"] == 0)
    for name in ("p", "q", "half", "material", "id", "kind", "module_class", "ring", "slot"):
        assert np.array_equal(f[name], b[name]), name
