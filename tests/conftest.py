import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import __graft_entry__ as ge  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests through the C ABI")


@pytest.fixture(scope="session")
def vp():
    return ge.load_package()


@pytest.fixture(autouse=True)
def _library_matches_sources(request):
    """GPU tests run only against a libvpbs.so built from the sources as they are now."""
    if request.node.get_closest_marker("gpu") is not None:
        ge.check_built()


@pytest.fixture(scope="session")
def O():
    return ge.load_oracle()


def rel_l2(a, b):
    import numpy as np

    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))
