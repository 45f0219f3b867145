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
    """One engine context for all GPU tests of the session (fails loudly without the HIP
    library).  Streams of at least 500 MBs per picture run on its second HIP stream, so
    batches of the golden cases use both lanes."""
    from h264_video_decoder_demo_amd import Engine
    eng = Engine(device=0, lane_split_mbs=500)
    yield eng
    eng.close()


_ENGINES = {}


@pytest.fixture(scope="session")
def engines():
    """Contexts per deblocking schedule (created on first use, closed at the end)."""
    yield _ENGINES
    for e in _ENGINES.values():
        e.close()


@pytest.fixture(params=["bands", "mb_wavefront", "mb_rows", "bands_small", "bands_single_row", "bands_two_lanes"])
def any_engine(request, engines):
    """Every intra / deblocking schedule; bands of 3 / 2 rows and of single rows (every row
    a band boundary); streams of 500 MBs and more on the second HIP stream."""
    from h264_video_decoder_demo_amd import Engine
    from h264_video_decoder_demo_amd import _lib
    kw = {"bands": dict(deblock_schedule=_lib.DEBLOCK_BANDS, intra_schedule=_lib.INTRA_BANDS),
          "mb_wavefront": dict(deblock_schedule=_lib.DEBLOCK_MB_WAVEFRONT, intra_schedule=_lib.INTRA_MB_WAVEFRONT),
          "mb_rows": dict(deblock_schedule=_lib.DEBLOCK_MB_ROWS, intra_schedule=_lib.INTRA_MB_WAVEFRONT),
          "bands_small": dict(deblock_schedule=_lib.DEBLOCK_BANDS, deblock_waves=3,
                              intra_schedule=_lib.INTRA_BANDS, intra_waves=2),
          "bands_single_row": dict(deblock_schedule=_lib.DEBLOCK_BANDS, deblock_waves=1,
                                   intra_schedule=_lib.INTRA_BANDS, intra_waves=1),
          "bands_two_lanes": dict(lane_split_mbs=500)}[request.param]
    if request.param not in engines:
        engines[request.param] = Engine(device=0, **kw)
    return engines[request.param]
