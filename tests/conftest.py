"""Test configuration: import paths and the `gpu` marker.

`-m "not gpu"` runs on any host (oracle vs golden vectors, host logic, library load and
exports); `-m gpu` needs an MI355X and runs the HIP path against the oracle.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lightning-lite-quantum-simulator_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP kernels)")
