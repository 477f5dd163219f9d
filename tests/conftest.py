import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

from __graft_entry__ import load_package  # noqa: E402

load_package()  # registers the package as ``rimu_amd`` (the native library loads lazily)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librimu_hip.so)")


@pytest.fixture(scope="session")
def rimu():
    return load_package()


@pytest.fixture(scope="session")
def gpu(rimu):
    if rimu.device_count() < 1:
        pytest.fail("GPU test run without a visible GPU")
    return rimu
