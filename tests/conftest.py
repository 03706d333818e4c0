import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "cuda-photon-mapper_amd"), os.path.join(REPO, "oracle"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through libpm.so)")


@pytest.fixture(scope="session")
def lib():
    import pmamd
    return pmamd.load()


@pytest.fixture(scope="session")
def ncpu():
    return max(1, min(16, len(os.sched_getaffinity(0))))
