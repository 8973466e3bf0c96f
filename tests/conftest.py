import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def gpu_pipeline():
    """Pipeline + stream on the GPU. Fails (does not skip) if the HIP library is missing."""
    from slosh_amd import pipeline as P
    if P.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    pl = P.MpmPipeline()
    st = P.Stream()
    yield pl, st
    st.close()
