import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "falling-dice_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU runs")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device presence only)
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def mmc():
    import mmc as m
    return m


@pytest.fixture(scope="session")
def po():
    import py_oracle
    return py_oracle


@pytest.fixture(params=["block", "graph"])
def loop_engine(request, monkeypatch):
    """Device-loop engine of the contexts the test creates (MMC_DEVICE_LOOP is
    read at context creation): the graph-replayed four-kernel loop ("graph",
    the default) or the workgroup-per-lattice solver ("block", opt-in, for
    lattices of up to 512 nodes)."""
    monkeypatch.setenv("MMC_DEVICE_LOOP", request.param)
    return request.param


@pytest.fixture(scope="session")
def gpu(mmc):
    """The GPU library; fails loudly (no fallback) when it is not built."""
    if not mmc.LIB_PATH.exists():
        pytest.fail(f"{mmc.LIB_PATH} missing: the HIP extension must be built before the GPU tests")
    mmc.lib()
    return mmc


def rel(a, b, floor=1e-300):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), floor))


def seeded_z(n_nodes, seed, t_sigma=0.02, r_sigma=0.02, squash=0.0, y=None, jr=0.0, ts=1.0):
    """A seeded configuration: Gaussian translations (units of the translation
    scale) and rotation vectors; squash > 0 moves nodes toward the lower platen
    by that fraction of their height above it."""
    rng = np.random.default_rng(seed)
    z = np.empty((n_nodes, 6))
    z[:, :3] = rng.normal(0, t_sigma, (n_nodes, 3))
    z[:, 3:] = rng.normal(0, r_sigma, (n_nodes, 3))
    if squash and y is not None:
        z[:, 1] -= squash * (y - jr) / ts
    return z.reshape(-1)
