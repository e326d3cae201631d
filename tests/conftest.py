import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def orc():
    """liboracle.so (built with the committed Makefile only when missing)."""
    from oracle import load_oracle

    return load_oracle()


@pytest.fixture(scope="session")
def bblib():
    """libbbgpu.so (built with bulbit_amd.build only when missing)."""
    from bulbit_amd import build, capi

    if not os.path.exists(capi.LIB_PATH):
        build.build()
    return capi.load_library()


def pixel_agreement(a, b, rel=1e-3, abs_floor=1e-6):
    """Fraction of pixels whose three channels agree within `rel` relative
    (|a - b| <= rel * max(|a|, |b|)) or within `abs_floor` absolute."""
    import numpy as np

    diff = np.abs(a - b)
    scale = np.maximum(np.abs(a), np.abs(b))
    ok = np.all((diff <= rel * scale) | (diff <= abs_floor), axis=-1)
    return float(ok.mean())
