import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def ctx():
    """One device context for the whole GPU session."""
    import ptwalker_amd as ptw

    c = ptw.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def host_ctx():
    """A context without a device: graph ingest only."""
    import ptwalker_amd as ptw

    c = ptw.Context(-1)
    yield c
    c.close()
