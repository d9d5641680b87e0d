import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref (reference tree present)")


@pytest.fixture(scope="session")
def ctx():
    import sfr3d_amd as S
    c = S.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ref():
    import oracle
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    r = oracle.ref()
    r.ref_init(1 << 31)
    return r
