"""pytest configuration: the `gpu` marker and shared fixtures.

GPU tests (`-m gpu`) call the HIP step through the C ABI of
libecmcrowd.so and compare it with the CPU oracle; they fail (not skip)
when the library or a gfx950 device is missing.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle_py
    oracle_py.lib()
    return oracle_py


@pytest.fixture(scope="session")
def native_gpu():
    from ecmgenerator_amd import _native
    L = _native.lib()
    assert L.ecm_device_available() == 1, "no gfx950 device visible to libecmcrowd.so"
    return L
