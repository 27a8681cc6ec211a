import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a HIP device is present; the HIP path must be the one that runs."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import cumf_sgd_amd  # noqa: F401  (raises if the library is not built)

    return 0
