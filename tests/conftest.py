import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libbonsai_gpu.so")


@pytest.fixture(scope="session")
def ctx():
    from bonsai_amd import Context

    c = Context(int(os.environ.get("BONSAI_DEVICE", "0")))
    yield c
    c.close()
