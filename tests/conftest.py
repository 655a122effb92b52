import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libtigerpm_hip.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    import tigerpm_amd as T
    ctx = T.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(autouse=True)
def _fresh_gpu_ctx_state(request):
    """Tests sharing the session context start from its defaults: default
    tuning, device order, no device budget (whatever an earlier test set)."""
    if "gpu_ctx" in request.fixturenames:
        import tigerpm_amd as T
        ctx = request.getfixturevalue("gpu_ctx")
        ctx.set_tuning()
        ctx.set_host_order(T.TPM_ORDER_DEVICE)
        ctx.set_budget(0)
    yield
