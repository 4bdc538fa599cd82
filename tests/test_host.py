"""CPU tests of the product's host side: the C-ABI library loads and exports
every entry point include/mmc.h declares, the struct layouts agree with the
header, the facade's lattice generators equal the oracle's bit for bit, the
simple-cubic family has the sizes of the benchmark workloads, and the CLI keeps
the reference's argument handling.  No compute call touches a GPU here."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "mmc.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(mmc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(mmc):
    lib = mmc.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(mmc.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}\b", out), n


def test_struct_layouts_match_header(mmc, tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text(f'#include "{HEADER}"\n#include <stdio.h>\n#include <stddef.h>\n'
                   "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(mmc_config),"
                   " sizeof(mmc_lattice), sizeof(mmc_objective_out), sizeof(mmc_opt_stats), sizeof(mmc_frame),"
                   " sizeof(mmc_state_stats), offsetof(mmc_config, seed));return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(mmc.Config), C.sizeof(mmc.Lattice), C.sizeof(mmc.ObjectiveOut), C.sizeof(mmc.OptStats),
            C.sizeof(mmc.Frame), C.sizeof(mmc.StateStats), mmc.Config.seed.offset]
    assert got == want


def test_defaults_match_header_defaults(mmc):
    c = mmc.Config()
    mmc.lib().mmc_config_defaults(C.byref(c))
    assert bytes(c) == bytes(mmc.default_config())


def test_last_error_never_null(mmc):
    assert mmc.lib().mmc_last_error(None) is not None
    assert mmc.lib().mmc_batch_last_error(None) is not None


def test_batch_rejects_bad_arguments_before_device_work(mmc):
    """Argument errors of the batched entry points come back as MMC_ERR_ARG
    without any device call (the host checks run first)."""
    lib = mmc.lib()
    h = C.c_void_p()
    assert lib.mmc_batch_create(None, 1, 0, C.byref(h)) == mmc.MMC_ERR_ARG
    cfgs = (mmc.Config * 1)(mmc.cubic_config(2, 2, 2))
    assert lib.mmc_batch_create(cfgs, 0, 0, C.byref(h)) == mmc.MMC_ERR_ARG
    assert lib.mmc_batch_info(None, 0, None, None, None, None, None) == mmc.MMC_ERR_ARG
    assert lib.mmc_batch_run(None, None, 0, None, None, 0) == mmc.MMC_ERR_ARG


@pytest.mark.parametrize("name", ["bcc_quick", "bcc_222", "cubic_8x8x4", "cubic_jitter", "cubic_seeded"])
def test_host_lattice_equals_oracle(mmc, po, name):
    cfg = {
        "bcc_quick": mmc.quick_config(),
        "bcc_222": mmc.default_config().replace(cells_x=2, cells_y=4, cells_z=2),
        "cubic_8x8x4": mmc.cubic_config(8, 8, 4),
        "cubic_jitter": mmc.cubic_config(5, 4, 3, sigma=1e-4, radius_mode=1, r_bot=3e-4, r_top=7e-4),
        "cubic_seeded": mmc.cubic_config(3, 3, 3).replace(seed=99),
    }[name]
    a = mmc.build_lattice(cfg)
    o = po.Solver(cfg, "orc").lattice()
    for k in ("x", "mass", "locked", "a", "b", "L0", "radius", "area", "inertia", "e0", "connected",
              "node_layer", "beam_layer", "beam_family"):
        assert np.array_equal(getattr(a, k), getattr(o, k)), k
    assert (a.width, a.height, a.depth) == (o.width, o.height, o.depth)


@pytest.mark.parametrize("cells,N,B", [((8, 8, 4), 405, 1044), ((32, 32, 16), 18513, 53328),
                                       ((64, 64, 32), 139425, 409760), ((6, 6, 6), 343, 882),
                                       ((16, 16, 16), 4913, 13872)])
def test_cubic_sizes(mmc, cells, N, B):
    a = mmc.build_lattice(mmc.cubic_config(*cells))
    assert len(a.mass) == N and len(a.L0) == B


def test_cubic_geometry(mmc):
    cfg = mmc.cubic_config(6, 5, 4)
    a = mmc.build_lattice(cfg)
    x = a.x.reshape(-1, 3)
    h = cfg.cell_size
    assert np.all(np.abs(a.L0 - h) < 4 * cfg.imperfection)
    e0 = a.e0.reshape(-1, 3)
    fam = a.beam_family
    for f in range(3):
        assert np.all(np.abs(e0[fam == f, f]) > 0.99)
    jr = cfg.joint_scale * cfg.max_radius
    mid_y = 0.5 * (x[a.a, 1] + x[a.b, 1])
    t = np.clip((mid_y - jr) / a.height, 0, 1)
    assert np.allclose(a.radius, cfg.min_radius * (1 - t) + cfg.max_radius * t, rtol=1e-15)
    jm = cfg.material.density * 4.0 / 3.0 * np.pi * jr ** 3 * 0.28
    total = cfg.material.density * (a.area * a.L0).sum() + len(a.mass) * jm
    assert a.mass.sum() == pytest.approx(total, rel=1e-12)
    locked = np.flatnonzero(a.locked)
    assert len(locked) == 3 and locked[0] % 6 == 0 and locked[1] % 6 == 2 and locked[2] % 6 == 2
    # top and bottom layers carry no imperfection: they meet the platens flat
    assert np.allclose(x[a.node_layer == 0, 1], jr) and np.allclose(x[a.node_layer == cfg.cells_y, 1], jr + a.height)


def test_cubic_gaussian_jitter(mmc):
    cfg = mmc.cubic_config(12, 10, 12, sigma=1e-4)
    a = mmc.build_lattice(cfg)
    x = a.x.reshape(-1, 3)
    nx, ny, nz = 12, 10, 12
    i = np.tile(np.arange(nx + 1), (ny + 1) * (nz + 1))
    dx = x[:, 0] - (-0.5 * a.width + i * cfg.cell_size)
    assert abs(dx.std() - 1e-4) < 1e-5 and abs(dx.mean()) < 1e-5
    lay = a.node_layer
    assert np.all(x[lay == 0, 1] == x[lay == 0, 1][0])


def cli():
    return ROOT / "falling-dice_amd" / "metamaterial_compression"


def test_cli_help_and_errors():
    r = subprocess.run([str(cli()), "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("metamaterial_compression --mode")
    r = subprocess.run([str(cli()), "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and r.stderr.strip() == "error: unknown option --bogus"
    r = subprocess.run([str(cli()), "--cells-x"], capture_output=True, text=True)
    assert r.returncode == 1 and r.stderr.strip() == "error: missing value for --cells-x"
    r = subprocess.run([str(cli()), "--mode", "render", "--output", "/tmp/mmc_cli_test"], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "render" in r.stderr


def test_device_construction_and_capture_reject_bad_arguments(mmc):
    """mmc_create_generated checks every configuration before it touches a
    device (cell counts, family), and the capture, lattice-readback and
    frame-observer entry points reject a missing context or batch."""
    lib = mmc.lib()
    h = C.c_void_p()
    bad = mmc.cubic_config(2, 2, 2).replace(cells_x=0)
    assert lib.mmc_create_generated(C.byref(bad), 1, 0, C.byref(h)) == mmc.MMC_ERR_ARG
    assert "cell counts must be positive" in lib.mmc_last_error(None).decode()
    fam = mmc.cubic_config(2, 2, 2).replace(lattice=7)
    assert lib.mmc_create_generated(C.byref(fam), 1, 0, C.byref(h)) == mmc.MMC_ERR_ARG
    assert "unknown lattice family" in lib.mmc_last_error(None).decode()
    good = mmc.cubic_config(2, 2, 2)
    assert lib.mmc_create_generated(C.byref(good), 0, 0, C.byref(h)) == mmc.MMC_ERR_ARG
    assert lib.mmc_create_generated(C.byref(good), 1, 5, C.byref(h)) == mmc.MMC_ERR_ARG
    assert lib.mmc_capture(None, None, None, None, None) == mmc.MMC_ERR_ARG
    assert lib.mmc_capture_mark(None) == mmc.MMC_ERR_ARG
    assert lib.mmc_lattice_info_get(None, 0, None) == mmc.MMC_ERR_ARG
    assert lib.mmc_batch_set_frame_observer(None, mmc.FRAME_OBSERVER(), None) == mmc.MMC_ERR_ARG
    assert lib.mmc_batch_evaluations(None) == 0
