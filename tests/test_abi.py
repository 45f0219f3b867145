"""The C ABI: every symbol include/h264g.h declares is exported by libh264g.so, the
ctypes mirrors have the C layouts, and calls without a device fail cleanly."""
import ctypes
import os
import re
import subprocess

import pytest

from h264_video_decoder_demo_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "h264g.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(h264g_\w+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("h264g_create", "h264g_picture_begin", "h264g_picture_submit", "h264g_batch_submit",
              "h264g_slot_readback", "h264g_sync", "h264g_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_every_declared_symbol():
    assert sorted(_lib._SIGNATURES) == declared_functions()


def _c_layout(tmp_path):
    """sizeof / offsetof of every ABI structure field, compiled from the header with gcc."""
    structs = {"h264g_mb": _lib.MbRecord, "h264g_pic_desc": _lib.PicDesc, "h264g_stream_desc": _lib.StreamDesc,
               "h264g_config": _lib.Config, "h264g_gen_params": _lib.GenParams}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "h264g.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, *_ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return dict(line.rsplit(" ", 1) for line in out.strip().splitlines()), structs


def test_ctypes_layouts_match_the_header(tmp_path):
    got, structs = _c_layout(tmp_path)
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, *_ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"
    assert _lib.MB_DTYPE.itemsize == int(got["h264g_mb"])
    for f in _lib.MB_DTYPE.names:
        assert _lib.MB_DTYPE.fields[f][1] == int(got[f"h264g_mb.{f}"]), f


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible")
def test_create_without_device_reports_nodev():
    from h264_video_decoder_demo_amd import Engine
    with pytest.raises(_lib.H264GError) as e:
        Engine(device=0)
    assert e.value.code == _lib.H264G_E_NODEV


def test_null_arguments_are_rejected():
    L = _lib.lib()
    assert L.h264g_create(None, None) == _lib.H264G_E_ARG
    assert L.h264g_sync(None) == _lib.H264G_E_ARG
    assert L.h264g_stream_open(None, None, None) == _lib.H264G_E_ARG
    assert b"NULL" in L.h264g_last_error() or L.h264g_last_error()


def test_generator_rejects_bad_parameters():
    from h264_video_decoder_demo_amd import Generator
    with pytest.raises(_lib.H264GError):
        Generator(1, qp_lo=40, qp_hi=20)
    with pytest.raises(_lib.H264GError):
        Generator(1, width_mbs=0)
