import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_PATH = os.path.join(ROOT, "tests", "golden", "golden.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libh264g.so")


@pytest.fixture(scope="session")
def golden():
    with open(GOLDEN_PATH) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def engine():
    """One engine context for all GPU tests of the session (fails loudly without the HIP library)."""
    from h264_video_decoder_demo_amd import Engine
    eng = Engine(device=0)
    yield eng
    eng.close()


@pytest.fixture(scope="session")
def engine_per_mb():
    """A second context that deblocks as a wavefront of single MBs."""
    from h264_video_decoder_demo_amd import Engine
    eng = Engine(device=0, deblock_rows=False)
    yield eng
    eng.close()


@pytest.fixture(params=["rows", "per_mb"])
def any_engine(request):
    """Both deblocking schedules (row-pipelined and per-MB wavefront)."""
    return request.getfixturevalue("engine" if request.param == "rows" else "engine_per_mb")
