import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def gpu():
    import phiray_amd as P

    g = P.Gpu.new(0)
    if g is None:
        pytest.fail("no HIP device: the gpu tests need one (run them on the GPU box)")
    yield g


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O

    O.lib()
    return O
