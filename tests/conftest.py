import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP core)")


@pytest.fixture(scope="session")
def gpu():
    """The HIP core, loaded, with a visible device (the core fails loudly if
    it is not built; no torch import, so only one HIP runtime is loaded)."""
    from lb_amd import abi

    abi.load_core()
    n, desc = abi.device_info()
    if n <= 0:
        pytest.fail(f"no HIP device visible ({desc})")
    print(f"[gpu] {desc}", flush=True)
    return desc
