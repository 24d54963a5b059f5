"""Shared pytest setup: import paths, the `gpu` marker and fixtures."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "aperturerendererv3-optix-geometry_amd")
for p in (PKG, os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def ap():
    import aperture

    aperture.lib()
    return aperture


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib

    return oracle_lib


@pytest.fixture(scope="session")
def gpu():
    import aperture

    if aperture.device_count() < 1:
        pytest.fail("a HIP device is required for -m gpu tests")
    aperture.check(aperture.lib().ap_device_set(0))
    print(f"[gpu] device 0: {aperture.device_name(0)}", flush=True)
    return aperture
