import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: larger meshes, still a few seconds")


def pytest_runtest_setup(item):
    # Name the case on stdout before any GPU work, so a failure that takes the process down
    # still says which test was running.
    if item.get_closest_marker("gpu") is not None:
        print(f"\n[gpu test] {item.nodeid}", flush=True)
