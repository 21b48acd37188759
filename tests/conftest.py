import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def _ensure_built():
    """Build what tests load only when nothing is built yet."""
    oracle_so = os.path.join(REPO, "oracle", "liboracle.so")
    if not os.path.exists(oracle_so):
        import build as oracle_build
        oracle_build.build_oracle()
    lib = os.path.join(REPO, "rtmuq2d_amd", "lib", "librtm_hip.so")
    if not os.path.exists(lib):
        from rtmuq2d_amd import build as pkg_build
        pkg_build.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu_available():
    import rtmuq2d_amd as R
    n = R.lib().rtm_device_count()
    if n < 1:
        pytest.fail("no HIP device visible: GPU tests must run on a GPU host")
    return n
