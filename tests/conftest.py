import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long-running statistical test")
    # GPU runs: start PyTorch's HIP runtime before the library's, so that the
    # device-hook tests can use torch tensors on the solver's buffers.
    expr = config.getoption("markexpr") or ""
    if "gpu" in expr and "not gpu" not in expr:
        try:
            import lattice_microbes_amd

            lattice_microbes_amd.torch_interop()
        except Exception:
            pass


@pytest.fixture(scope="session")
def gpu():
    import lattice_microbes_amd as lmx

    if lmx.device_count() < 1:
        pytest.fail("no HIP device visible: the gpu tests need a machine with a HIP device")
    return 0
