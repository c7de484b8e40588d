import os
import sys

import pytest

# HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default, and
# the variable may already be set to that).  The GPU session holds several
# contexts, each with up to three streams, and the serving lanes of one
# context must not share a hardware queue or their batches run one after the
# other; this takes effect before HIP starts.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

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


def _graph_context(scale, weights):
    import ptwalker_amd as ptw
    from ptwalker_amd.workloads import RMAT_ABC

    print(f"\n[fixture] R-MAT scale {scale} on its own context", flush=True)
    c = ptw.Context(0)
    a, b, cc = RMAT_ABC
    gi = c.rmat(scale=scale, edge_factor=16, a=a, b=b, c=cc, seed=1)
    if weights:
        c.random_weights(seed=1)
    rp, col = c.export_csr()
    return c, gi, rp, col


@pytest.fixture(scope="session")
def ctx22():
    """R-MAT scale 22 (the single-source and serving configurations), resident
    on a context of its own for the whole GPU session."""
    c, gi, rp, col = _graph_context(22, False)
    yield c, gi, rp, col
    c.close()


@pytest.fixture(scope="session")
def ctx24():
    """R-MAT scale 24 with U(0, 1] weights (the DeepWalk corpus and batched
    configurations), resident on a context of its own."""
    c, gi, rp, col = _graph_context(24, True)
    yield c, gi, rp, col
    c.close()
