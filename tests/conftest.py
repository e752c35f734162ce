import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU test")


def _ensure_built():
    lib = ROOT / "e2e_gpu_dp_amd" / "libe2edp.so"
    if not lib.exists():
        from e2e_gpu_dp_amd.build import build
        build(verbose=False)
    import oracle
    if not oracle.SO.exists():
        oracle.compile_lib()


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    # torch brings its own HIP runtime; it must initialise before libe2edp binds one
    import torch
    torch.cuda.is_available()
    import e2e_gpu_dp_amd as e2e
    if e2e.device_count() < 1:
        pytest.fail("no HIP device visible (gpu tests must run on the GPU box)")
    return True
