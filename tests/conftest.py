import gzip
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "gpu-svo_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def bunny_path(tmp_path_factory):
    """The reference's sample mesh (bunny.obj), from the committed fixture."""
    d = tmp_path_factory.mktemp("bunny")
    p = d / "bunny.obj"
    with gzip.open(os.path.join(GOLDEN, "bunny.obj.gz"), "rb") as f:
        p.write_bytes(f.read())
    return str(p)


@pytest.fixture(scope="session")
def bunny(bunny_path):
    import pyoracle as po
    st, P, T = po.load_obj(bunny_path)
    assert st == 0
    return P, T


@pytest.fixture(scope="session")
def gpu_ctx():
    import svo_amd
    ctx = svo_amd.Context(0)
    yield ctx
    ctx.close()
