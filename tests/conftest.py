import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "mixed-mlx_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def runtime():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from msq import capi
    rt = capi.Runtime(0)
    yield rt
    rt.close()


def pytest_report_header(config):
    return f"msq root: {ROOT}"
