import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "webgpu-fluid-sim_amd")
for p in (ROOT, PKG_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnbody_sim.so)")
    config.addinivalue_line("markers", "slow: large sizes")


def pytest_runtest_logstart(nodeid, location):
    # Name the case before any GPU work so a fault can be attributed.
    print(f"\n[case] {nodeid}", flush=True)
