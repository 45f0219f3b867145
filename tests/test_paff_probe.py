"""Field pictures (PAFF): what the reference decoder does with them.

oracle/paff_probe.cpp writes a small Annex-B stream (eight frames of the 8-frame,
2-B-frame GOP, then a top/bottom field pair, then a P frame) and decodes it with the
reference's own decoder built with AddressSanitizer (make -C oracle paff).  These
tests pin the two findings the engine's refusal of field pictures rests on: every P
field is a NULL dereference in the reference's field reference-list initialisation
(H264RefPicList.cpp:1197-1230), and I fields land in DPB entries that held frames
(H264PictureBase.cpp:668-687) with half-height frame views (H264PictureBase.cpp:
205-208).  They need /root/reference and g++, so they are skipped where either is
missing; nothing here touches a GPU.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
PROBE = os.path.join(ORACLE, "_ref", "paff_probe")
REF = "/root/reference/h264_video_decoder_demo"

pytestmark = pytest.mark.skipif(not os.path.isdir(REF) or shutil.which("g++") is None,
                                reason="needs the reference sources and g++")


@pytest.fixture(scope="module")
def probe():
    subprocess.run(["make", "-s", "-j8", "-C", ORACLE, "paff"], check=True, capture_output=True, timeout=600)
    return PROBE


def run_probe(probe, tmp_path, w, h, reorder, fields):
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=0:detect_leaks=0")
    r = subprocess.run([probe, str(w), str(h), str(reorder), fields, str(tmp_path / "s.h264")],
                       capture_output=True, text=True, timeout=300, env=env)
    return r, r.stdout + r.stderr


def table(out):
    rows = []
    lines = out.splitlines()
    start = next(i for i, l in enumerate(lines) if l.startswith("PAFF_PROBE"))
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
    assert "PAFF_PROBE" not in out  # the decoder never returned


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
