"""Lattices built on the device (mmc_create_generated) against the host
construction (mmc_lattice_build), which equals the reference's
Simulator::build + distributeMass (MC/src/sim.cpp:49-135) bit for bit
(tests/test_host.py, tests/test_oracle_pin.py).

Exact: topology (beam ends), node and beam layers, beam families, the locked
DOFs, the self-contact exclusion adjacency, and everything the construction
makes of sums, products, quotients and square roots: node positions of the BCC
family and of the unjittered cubic family, rest lengths, directors, radii,
areas, lumped masses and the total mass.  Within one unit in the last place:
the second moments pi r^4 / 4 (the device rounds r^4 correctly; glibc's pow,
which the host uses, misses by one unit about once in a thousand: measured
185 of 200,000 radii against exact rational powers).  Within 16 units in the
last place of the lattice's extent: the Gaussian node jitter of the cubic family (Box-Muller through the
device's log, sin and cos), and what follows from those positions to 1e-14.
A batch built on the device then runs the schedule of the one built on the
host: bit for bit when the lattices are, and within the tolerance band of the
gradient tolerance otherwise (1e-6 of the peak force at tolerance 1e-7)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases(mmc):
    return {
        "bcc_specimen": mmc.default_config().replace(cells_x=2, cells_y=4, cells_z=2),
        "bcc_343_two_segments": mmc.default_config().replace(cells_x=3, cells_y=4, cells_z=3, elements_per_strut=2,
                                                              seed=12345),
        "bcc_232_five_segments": mmc.default_config().replace(cells_x=2, cells_y=3, cells_z=2, elements_per_strut=5,
                                                               seed=7),
        "cubic_8x8x4": mmc.cubic_config(8, 8, 4),
        "cubic_jitter_beam_radii": mmc.cubic_config(5, 4, 3, sigma=1e-4, radius_mode=1, r_bot=3e-4, r_top=7e-4),
        "sweep_design": mmc.batch_workload("sweep", 0, indices=[5])[0],
        "uneven_lattice": mmc.batch_workload("uneven", 0, indices=[4])[0],
    }


def _adjacency(host, N):
    nb = [[] for _ in range(N)]
    for key in host.connected.tolist():
        u, v = key >> 32, key & 0xFFFFFFFF
        nb[u].append(v)
        nb[v].append(u)
    off = np.zeros(N + 1, np.int64)
    lists = []
    for i in range(N):
        s = sorted(set(nb[i]))
        off[i + 1] = off[i] + len(s)
        lists.extend(s)
    return off, np.array(lists, np.int64)


@pytest.mark.parametrize("case", ["bcc_specimen", "bcc_343_two_segments", "bcc_232_five_segments", "cubic_8x8x4",
                                  "cubic_jitter_beam_radii", "sweep_design", "uneven_lattice"])
def test_generated_lattice_matches_host(gpu, po, case):
    mmc = gpu
    cfg = _cases(mmc)[case]
    host = mmc.build_lattice(cfg)
    ctx = mmc.Context.generated(cfg)
    dev, adj_off, adj = ctx.read_lattice()
    info = ctx.lattice_info()
    N, B = len(host.mass), len(host.L0)
    assert (info.n_nodes, info.n_beams, info.n_exclusions) == (N, B, len(host.connected))
    for name in ("a", "b", "node_layer", "beam_layer", "beam_family", "locked"):
        assert np.array_equal(getattr(dev, name), getattr(host, name)), name
    off, lists = _adjacency(host, N)
    assert np.array_equal(adj_off, off) and np.array_equal(adj, lists)
    assert (info.width, info.height, info.depth) == (host.width, host.height, host.depth)
    oracle_mass = po.Solver(cfg, "orc").mass  # distributeMass's total, the reference's order
    jitter = cfg.lattice == mmc.MMC_LATTICE_CUBIC and cfg.position_sigma > 0
    if jitter:
        ext = float(np.abs(host.x).max())
        dx = float(np.abs(dev.x - host.x).max())
        print(case, "largest position difference", dx, "=", dx / (np.finfo(float).eps * ext), "ulp of the extent")
        assert dx <= 16 * np.finfo(float).eps * ext
        for name in ("L0", "e0", "radius", "area", "mass"):
            d, h = getattr(dev, name), getattr(host, name)
            assert float(np.abs(d - h).max()) <= 1e-14 * float(np.abs(h).max()), name
        assert abs(info.mass - oracle_mass) <= 1e-14 * oracle_mass
    else:
        for name in ("x", "L0", "e0", "radius", "area", "mass"):
            assert np.array_equal(getattr(dev, name), getattr(host, name)), name
        assert info.mass == oracle_mass
    # the second moment: the device's pow(r, 4) against glibc's
    ulps = np.abs(dev.inertia - host.inertia) / np.spacing(host.inertia)
    print(case, "second moments off by one ulp:", int(np.sum(ulps > 0)), "of", B)
    assert float(ulps.max()) <= 1.0 and int(np.sum(ulps > 0)) <= max(2, B // 100)


def test_generated_batch_runs_the_host_batch(gpu):
    """Three lattices in parity form (gradient tolerance 1e-7), built on the
    device and on the host, through their whole schedules."""
    mmc = gpu
    cfgs = [mmc.tight(mmc.cubic_config(3, 4, 3, load=6, unload=3, sigma=1e-4, radius_mode=1, r_bot=3e-4, r_top=7e-4,
                                       max_iterations=4000)),
            mmc.tight(mmc.cubic_config(4, 4, 3, load=8, unload=4, max_iterations=4000)),
            mmc.tight(mmc.batch_workload("sweep", 0, load=8, unload=4, indices=[2])[0])]
    fd, sd, ed = mmc.BatchSimulator(cfgs, mmc.MMC_FP64, device_lattices=True).run()
    fh, sh, eh = mmc.BatchSimulator(cfgs, mmc.MMC_FP64).run()
    assert sd == sh == [0] * len(cfgs), (ed, eh)
    for b, cfg in enumerate(cfgs):
        host = mmc.build_lattice(cfg)
        dev, _, _ = mmc.Context.generated(cfg).read_lattice()
        same = all(np.array_equal(getattr(dev, k), getattr(host, k)) for k in ("x", "inertia", "radius"))
        assert len(fd[b]) == len(fh[b])
        peak = float(fh[b]["reaction_force"].max())
        err = float(np.abs(fd[b]["reaction_force"] - fh[b]["reaction_force"]).max()) / peak
        print(b, "lattices identical:", same, "force difference of peak", err)
        if same:
            assert fd[b].tobytes() == fh[b].tobytes()
        assert err <= 1e-6
