"""Shared test setup: the `gpu` marker, in-tree builds of missing libraries."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def _ensure_built():
    # The oracle is a few hundred lines of C: build it if it is missing.
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    if not os.path.exists(os.path.join(ROOT, "fasttransforms_amd", "libfasttransforms_amd.so")):
        subprocess.run([sys.executable, os.path.join(ROOT, "fasttransforms_amd", "build.py")], check=True,
                       stdout=subprocess.DEVNULL)


_ensure_built()


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ft():
    import fasttransforms_amd
    return fasttransforms_amd


@pytest.fixture(scope="session")
def O():
    import oracle
    return oracle
