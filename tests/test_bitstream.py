"""The bitstream front end (h264g_parser_*) on a prefix of the reference's own CAVLC
sample stream (tests/golden/heavyhand_cavlc_15.*, written by make_real.py from the
reference decoder's output): parse on the CPU, reconstruct on the GPU, and compare
every output picture with the reference decoder's."""
import hashlib
import json
import os

import numpy as np
import pytest

from h264_video_decoder_demo_amd import _lib
from h264_video_decoder_demo_amd.bitstream import Bitstream
from h264_video_decoder_demo_amd.decoder import H264_DECODE_ERROR_CODE_FILE_END, H264VideoDecoder

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STREAM = os.path.join(HERE, "heavyhand_cavlc_15.h264")
REF = json.load(open(os.path.join(HERE, "heavyhand_cavlc_15.json")))["pictures"]


def _decode_order(ref):
    return [p for p in sorted(ref, key=lambda p: p["decode_index"])]


def test_parser_follows_the_reference_picture_order():
    """Every slice parses to its end (the parser refuses a slice whose MBs do not cover
    the picture or whose data runs out), in the reference's decoding order, with the
    reference's PicOrderCnt and slice type for each picture."""
    bs = Bitstream(STREAM)
    assert (bs.width_mbs, bs.height_mbs) == (120, 68)
    pics = list(bs.pictures())
    want = _decode_order(REF)
    assert [(p.desc.poc, p.desc.slice_type) for p in pics] == [(w["poc"], w["slice_type"]) for w in want]
    assert pics[0].idr and not any(p.idr for p in pics[1:])
    # reference lists name slots of earlier reference pictures only, never the destination
    for p in pics:
        for l in range(2):
            refs = [p.desc.ref_slot[l][i] for i in range(p.desc.num_ref_idx_active[l])]
            assert all(0 <= r < bs.dpb_slots and r != p.desc.dst_slot for r in refs), (p.desc.poc, refs)
    # a second pass from the start gives the same pictures
    again = list(Bitstream(STREAM).pictures())
    for a, b in zip(pics, again):
        assert np.array_equal(a.mbs, b.mbs) and np.array_equal(a.levels, b.levels)


def test_parser_refuses_cabac_and_damaged_streams():
    """A CABAC PPS is refused by name; damaged slice data ends in an error or a parse
    that still stays inside its buffers (the library never reads past the stream)."""
    data = bytearray(open(STREAM, "rb").read())
    # entropy_coding_mode_flag is the third bit of the PPS payload (pps_id, sps_id: ue 0)
    k = data.find(b"\x00\x00\x01\x68")
    data[k + 4] |= 0x20
    with pytest.raises(_lib.H264GError, match="CABAC"):
        list(Bitstream(bytes(data)).pictures())
    rng = np.random.default_rng(5)
    clean = open(STREAM, "rb").read()
    first_slice = clean.find(b"\x00\x00\x01\x65")
    for _ in range(12):
        d = bytearray(clean)
        for pos in rng.integers(first_slice + 16, len(d), 4):
            d[pos] ^= int(rng.integers(1, 256))
        try:
            for _p in Bitstream(bytes(d)).pictures():
                pass
        except (_lib.H264GError, ValueError):
            pass
    with pytest.raises(ValueError):
        Bitstream(b"\x00\x00\x01\x09\xf0")  # an access unit delimiter, no picture


@pytest.mark.gpu
def test_real_stream_equals_reference_decoder(engine):
    """The 15 pictures decoded through H264VideoDecoder (host parse, GPU reconstruction)
    are handed to the callback in the reference decoder's output order, and each one's
    planes equal the reference decoder's byte for byte."""
    got = []

    def callback(pic, user, err):
        if pic is None:
            assert err == H264_DECODE_ERROR_CODE_FILE_END
            return 0
        got.append((pic.PicOrderCnt, hashlib.sha256(pic.m_pic_buff_luma.tobytes()).hexdigest(),
                    hashlib.sha256(pic.m_pic_buff_cb.tobytes()).hexdigest(),
                    hashlib.sha256(pic.m_pic_buff_cr.tobytes()).hexdigest()))
        return 0

    dec = H264VideoDecoder(engine)
    dec.set_output_frame_callback_functuin(callback, None)
    assert dec.open(STREAM) == 0
    assert [g[0] for g in got] == [r["poc"] for r in REF]
    for g, r in zip(got, REF):
        assert g[1:] == (r["y"], r["cb"], r["cr"]), f"POC {r['poc']} (decoding index {r['decode_index']})"
