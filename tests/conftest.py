"""pytest configuration: import paths and the `gpu` marker.

Tests marked `gpu` need an MI355X (HIP device); everything else runs on the
CPU.  The CPU oracle (oracle/) is built on first use.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cf-3dgs_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")
