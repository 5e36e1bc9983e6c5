import os
import tempfile
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    # launch-shape tuning cache of the test session (not the user's)
    os.environ.setdefault("KUBOMSD_TUNING_CACHE", os.path.join(tempfile.mkdtemp(prefix="kmsd_tuning_"),
                                                               "tuning.json"))
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP) device; never skipped silently")


@pytest.fixture(scope="session")
def hip_ctx():
    """One HIP context for the whole GPU session; fails (not skips) without a device."""
    import torch
    from kubomsd_amd.backends.hip import HipContext
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but torch.cuda.is_available() is False")
    return HipContext(0)
