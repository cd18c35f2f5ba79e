import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    from gammaray_amd.api import Context
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    torch.cuda.init()
    ctx = Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def host_ctx():
    from gammaray_amd.api import Context, GR_HOST_ONLY
    ctx = Context(GR_HOST_ONLY)
    yield ctx
    ctx.close()
