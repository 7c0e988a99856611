import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libnntc_ref.so")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "nntc_encode")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the gfx950 arm)")


@pytest.fixture(scope="session")
def ref_arm():
    from nntc_amd import abi

    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref is not built (no /root/reference where this ran)")
    arm = abi.Arm(REF_LIB, "nntc_ref_")
    arm.lib.nntc_ref_set_threads(min(16, os.cpu_count() or 1))
    return arm


@pytest.fixture(scope="session")
def hip_arm():
    from nntc_amd import abi

    return abi.Arm()
