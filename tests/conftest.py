import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built librsr_amd.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    from rsr_amd import api
    ctx = api.Context(0)
    # The library defaults to reference-order summation (pinned by
    # test_gpu_eval.py::test_default_summation_is_reference_order); the shared context starts
    # in Kahan mode, the ambient mode of the tests that do not pick one themselves.
    ctx.set_summation(False)
    yield ctx
    ctx.close()
