import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: large workloads")


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    from ktplace_amd.simpl import Context

    def make(design=None):
        ctx = Context(0)
        if design is not None:
            ctx.upload(design)
        return ctx

    return make
