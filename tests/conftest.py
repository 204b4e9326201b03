import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "xpbi-webgpu_amd"), os.path.join(REPO, "oracle"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (oracle/).  XPBI_ORACLE_LIB names another build of it: the
    mutation test runs the oracle's pins against copies with planted mistakes."""
    import oracle_py
    alt = os.environ.get("XPBI_ORACLE_LIB")
    if alt:
        oracle_py.use_library(alt).__enter__()
    else:
        oracle_py.build()
    return oracle_py
