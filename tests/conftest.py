import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "boostfractor.jl_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture
def engine_options():
    """set engine options of the thread's context (bf_ctx_set_option) for one
    test; the previous values are restored at teardown"""
    import BoostFractor as B
    ctx = B.context()
    saved = {}

    def set_(**kw):
        for k, v in kw.items():
            saved.setdefault(k, ctx.get_option(k))
            ctx.set_option(k, v)

    yield set_
    for k, v in saved.items():
        ctx.set_option(k, v)
