import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "liboracle_j2k.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return lib


@pytest.fixture(scope="session")
def oracle():
    _ensure_oracle()
    from tests import oracle_ref
    return oracle_ref


@pytest.fixture(scope="session")
def session():
    import signinum_amd as S
    with S.HipSession() as s:
        yield s
