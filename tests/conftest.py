import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def ctx():
    """One library context on cuda:0 bound to torch's current stream."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from varrcwa_amd._lib import Context
    c = Context(0)
    yield c
    c.close()
