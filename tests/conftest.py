import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
for p in (str(REPO), str(REPO / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built liblgdhip.so")


@pytest.fixture(scope="session")
def hip():
    """The liblgdhip binding with device 0 selected.  Fails (does not skip)
    when the library or the GPU is missing: GPU tests must run native code."""
    from libgeodecomp_amd import lgdhip

    lgdhip.load()
    n = lgdhip.device_count()
    assert n >= 1, "no HIP device visible"
    lgdhip.call("lgdhip_set_device", 0)
    print(f"[gpu] {lgdhip.device_name(0)}", flush=True)
    return lgdhip


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib

    oracle_lib.load()
    return oracle_lib
