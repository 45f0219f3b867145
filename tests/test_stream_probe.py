"""Field pictures (PAFF) and temporal direct prediction: what the reference decoder does.

oracle/stream_probe.cpp writes a small Annex-B stream (eight frames of the 8-frame,
2-B-frame GOP, then optionally a top/bottom field pair, then a P frame) and decodes it
with the reference's own decoder built with AddressSanitizer (make -C oracle probe).
These tests pin the findings the engine's refusals rest on: every P field is a NULL
dereference in the reference's field reference-list initialisation
(H264RefPicList.cpp:1197-1230); I fields land in DPB entries that held frames
(H264PictureBase.cpp:668-687) with half-height frame views (H264PictureBase.cpp:
205-208); and temporal direct prediction looks up the co-located picture itself in
RefPicList0 (H264InterPrediction.cpp:1465-1478), so that it either runs off the list
(NULL dereference) or divides by td = 0 (H264InterPrediction.cpp:1643).  They need
/root/reference and g++, so they are skipped where either is missing; nothing here
touches a GPU.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
PROBE = os.path.join(ORACLE, "_ref", "stream_probe")
REF = "/root/reference/h264_video_decoder_demo"

pytestmark = pytest.mark.skipif(not os.path.isdir(REF) or shutil.which("g++") is None,
                                reason="needs the reference sources and g++")


@pytest.fixture(scope="module")
def probe():
    subprocess.run(["make", "-s", "-j8", "-C", ORACLE, "probe"], check=True, capture_output=True, timeout=600)
    return PROBE


def run_probe(probe, tmp_path, w, h, reorder, fields, temporal=0):
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=0:detect_leaks=0")
    r = subprocess.run([probe, str(w), str(h), str(reorder), fields, str(tmp_path / "s.h264"), str(temporal)],
                       capture_output=True, text=True, timeout=300, env=env)
    return r, r.stdout + r.stderr


def table(out):
    rows = []
    lines = out.splitlines()
    start = next(i for i, l in enumerate(lines) if l.startswith("STREAM_PROBE"))
    for l in lines[start + 2:]:
        f = l.split()
        if len(f) != 9 or not f[0].isdigit():
            break
        rows.append(dict(num=int(f[0]), type=f[1], entry=int(f[2]), reused=f[3] == "yes", rows=int(f[4]),
                         buf=int(f[5])))
    return rows


def test_p_fields_crash_the_reference(probe, tmp_path):
    r, out = run_probe(probe, tmp_path, 4, 4, 0, "P")
    assert r.returncode != 0
    assert "SEGV" in out
    assert "Initialisation_process_for_reference_picture_lists_in_fields" in out
    assert "STREAM_PROBE" not in out  # the decoder never returned


def test_i_fields_land_in_frame_entries_with_half_height_views(probe, tmp_path):
    r, out = run_probe(probe, tmp_path, 4, 4, 0, "I")
    rows = table(out)
    assert [x["type"] for x in rows] == ["frame"] * 8 + ["field"] * 2 + ["frame"]
    fields = [x for x in rows if x["type"] == "field"]
    assert all(x["reused"] for x in fields)            # both fields went into entries that held frames
    assert fields[0]["entry"] != fields[1]["entry"]    # one entry per field
    assert all(x["rows"] == 4 * 16 // 2 for x in fields)   # the frame view is half height
    assert all(x["buf"] == 4 * 4 * 384 for x in fields)    # behind it, the full frame buffer


def test_i_fields_in_never_used_entries_overflow(probe, tmp_path):
    # no VUI: the reorder depth is inferred from the level (16 at 4x4 MBs), so the
    # field pair goes to entries never used before, allocated at half a frame
    r, out = run_probe(probe, tmp_path, 4, 4, -1, "I")
    rows = table(out)
    fields = [x for x in rows if x["type"] == "field"]
    assert fields and not any(x["reused"] for x in fields)
    assert all(x["buf"] == 4 * 4 * 384 // 2 for x in fields)
    assert "WRITE of size" in out and "Sample_construction_process_for_I_PCM_macroblocks" in out


def test_spatial_direct_b_pictures_decode(probe, tmp_path):
    r, out = run_probe(probe, tmp_path, 4, 4, 0, "N", temporal=0)
    rows = table(out)
    assert len(rows) == 9 and "DEADLYSIGNAL" not in out


@pytest.mark.parametrize("temporal,what", [(1, "SEGV"), (2, "FPE")])
def test_temporal_direct_crashes_the_reference(probe, tmp_path, temporal, what):
    # one active reference per list: the co-located picture is not in RefPicList0 and the
    # search runs onto a NULL entry; two: it is found, pic0 == pic1 and td = 0
    r, out = run_probe(probe, tmp_path, 4, 4, 0, "N", temporal=temporal)
    assert r.returncode != 0 and "STREAM_PROBE" not in out
    assert f"AddressSanitizer: {what}" in out
    assert "Derivation_process_for_temporal_direct_luma_motion_vector_and_reference_index_prediction_mode" in out
