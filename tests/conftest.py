import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libdemsph.so on the GPU)")
    # The oracle is the checker; it is plain C and builds in about a second.
    lib = ROOT / "oracle" / "liboracle.so"
    srcs = list((ROOT / "oracle").glob("*.[ch]"))
    if not lib.exists() or any(s.stat().st_mtime > lib.stat().st_mtime for s in srcs):
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


def pytest_collection_modifyitems(config, items):
    # Print the case name before each GPU test so that a fault names it.
    pass


@pytest.fixture(autouse=True)
def _announce(request):
    if request.node.get_closest_marker("gpu") is not None:
        print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
