import importlib
import os
import sys
import tarfile

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def cms():
    return importlib.import_module("coordinated-meta-storms_amd")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


def has_gpu() -> bool:
    """Device check through libcms itself. torch is deliberately not imported
    in processes that load libcms: it bundles its own HIP runtime, and two HIP
    runtimes in one process do not share the device."""
    try:
        cms = importlib.import_module("coordinated-meta-storms_amd")
        return cms.device_count() > 0
    except Exception:
        return False


class Golden:
    """An unpacked golden case: the reference's inputs and outputs."""

    def __init__(self, case, tmp):
        d = os.path.join(tmp, case)
        if not os.path.exists(d):
            os.makedirs(d)
            with tarfile.open(os.path.join(GOLDEN, case + ".tar.gz")) as tf:
                tf.extractall(d, filter="data")
        self.dir = d
        self.root = os.path.join(d, "input")
        self.table = os.path.join(self.root, "table.txt")

    def read(self, name) -> bytes:
        with open(os.path.join(self.dir, name), "rb") as f:
            return f.read()

    def hex_sims(self, cp=True):
        """(i, j, sim) of the reference Calc_sim for every pair, exact floats."""
        rows = []
        for line in self.read("hex_rT.txt" if cp else "hex_rF.txt").decode().split("\n"):
            if line.strip():
                i, j, h = line.split()
                rows.append((int(i), int(j), float.fromhex(h)))
        i = np.array([r[0] for r in rows], np.int32)
        j = np.array([r[1] for r in rows], np.int32)
        s = np.array([r[2] for r in rows], np.float32)
        return i, j, s


GOLDEN_CASES = ["toy", "quirks", "config0", "core"]


@pytest.fixture(scope="session")
def golden(tmp_path_factory):
    base = str(tmp_path_factory.mktemp("golden"))
    return lambda case: Golden(case, base)


@pytest.fixture(scope="session")
def gpu_ok():
    if not has_gpu():
        pytest.skip("no HIP device")
    return True
