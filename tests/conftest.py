import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch.device("cuda:0")


@pytest.fixture(params=["mfma", "valu"])
def kernel_family(request, monkeypatch):
    """Runs a test once on the matrix-core kernels (the default) and once on
    the direct FP32 vector kernels; the library reads the switches at every
    call."""
    if request.param == "valu":
        monkeypatch.setenv("TUCU_GAUSS_PATH", "valu")
        monkeypatch.setenv("TUCU_FIR_PATH", "valu")
    else:
        monkeypatch.delenv("TUCU_GAUSS_PATH", raising=False)
        monkeypatch.delenv("TUCU_FIR_PATH", raising=False)
    return request.param
