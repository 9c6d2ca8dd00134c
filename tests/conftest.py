import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "amt-bh-benchmark_amd")
for p in (PKG, REPO, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle.Oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle

    if not oracle.reference_available():
        pytest.skip("oracle/_ref (reference-compiled checker) not built here")
    return oracle.Reference()


@pytest.fixture(scope="session")
def gpu():
    from bhgpu import lib as bl

    ctx = bl.Context(0)
    yield ctx
    ctx.close()
