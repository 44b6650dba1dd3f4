import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels through libmcgb.so (needs an MI355X)")
    config.addinivalue_line("markers", "slow: statistical test that takes more than ~20 s")
