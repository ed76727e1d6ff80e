import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the GPU)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from sod2d_refmap_amd.device import init
    init(0)
    return torch.device("cuda:0")


@pytest.fixture(autouse=True)
def _library_defaults_after_gpu_test(request):
    """Process-global library settings (numerical parameters, stage
    configuration mode, rk_4_main implementation) go back to their defaults
    after every GPU test, also when it fails, so that no test leaks its
    settings into the next."""
    yield
    if request.node.get_closest_marker("gpu") is None:
        return
    from sod2d_refmap_amd import device
    if not device._INITIALISED:
        return
    from sod2d_refmap_amd import _lib, cases
    device.set_numerical_params(device.numerical_params(Rgas_inf=cases.GasParams().Rgas))
    _lib.call("sod_set_autotune", 0)
    _lib.call("sod_set_rk4_path", 0)
