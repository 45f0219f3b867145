"""Host motion-vector derivation (h264g_mvfield_*) against the reference's own.

The fixtures under tests/golden/mvd_*.npz hold, for generated streams, the mvd of every
(sub-)partition that the reference decoder's prediction (H264InterPrediction.cpp:
1660-2001, driven by oracle/ref_harness.cpp) needs to reproduce the generator's final
motion vectors; the harness decoded each stream again from those mvds and found every
vector equal, and in the *_direct streams every reference index and vector that the
reference derived itself for B_Skip, B_Direct_16x16 and B_Direct_8x8 (spatial direct,
H264InterPrediction.cpp:892-1432) equal to the generator's.  Deriving the vectors from
the same mvds must give the generator's vectors back, partition by partition, P_Skip
and direct MBs included.  No GPU is used.
"""
import glob
import json
import os

import numpy as np
import pytest

from h264_video_decoder_demo_amd import Generator, MotionField, _lib

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(glob.glob(os.path.join(GOLDEN, "mvd_*.npz")))
INTER = (_lib.MB_P, _lib.MB_P_SKIP, _lib.MB_B)


def _quadrant(blk):
    return (blk // 8) * 2 + (blk % 4) // 2


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[4:-4] for p in CASES])
def test_derived_mvs_equal_generator(path):
    z = np.load(path, allow_pickle=False)
    g = Generator(int(z["config"]), **json.loads(str(z["overrides"])))
    assert g.digest == int(z["digest"]), "generator changed: regenerate with tests/golden/make_mvd.py"
    mvd = z["mvd"]
    f = MotionField(g.width_mbs, g.height_mbs)
    checked = {4: 0, 5: 0, 6: 0}
    slot_pic = {}  # DPB slot -> the records last decoded into it
    direct = 0
    for i in range(g.num_pictures):
        desc, mbs, _ = g.picture(i)
        col = slot_pic[desc.ref_slot[1][0]] if desc.slice_type == _lib.SLICE_B else None
        inp = mbs.copy()
        dq = (inp["flags"][:, None] >> (4 + np.arange(4))) & 1  # direct quadrants: indices are derived
        inp["ref_idx"][:, 0][dq == 1] = 99
        inp["ref_idx"][:, 1][dq == 1] = 99
        got = f.derive_picture(inp, mvd[i], col=col)
        slot_pic[desc.dst_slot] = got
        assert (got["ref_idx"] == mbs["ref_idx"]).all(), i
        direct += int(dq.sum())
        for a in np.nonzero(np.isin(mbs["mb_type"], INTER))[0]:
            for blk in range(16):
                q = _quadrant(blk)
                for l in range(2):
                    if mbs["ref_idx"][a, l, q] < 0:
                        continue
                    assert tuple(got["mv"][a, l, blk]) == tuple(mbs["mv"][a, l, blk]), \
                        (i, int(a), int(mbs["mb_type"][a]), l, blk)
            checked[int(mbs["mb_type"][a])] += 1
    assert checked[4] > 0 and checked[5] + checked[6] > 0, checked
    if json.loads(str(z["overrides"])).get("direct_permille"):
        assert direct > 0


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[4:-4] for p in CASES])
def test_derived_intra_modes_equal_generator(path):
    """The intra modes of I_NxN MBs derived from the reference-solved prev / rem syntax
    (8.3.1.1 / 8.3.2.1, getIntra4x4PredMode / getIntra8x8PredMode) are the generator's,
    I4x4 and I8x8 MBs beside every kind of neighbour, and with constrained intra
    prediction (inter neighbours predict DC) in the *_cip streams."""
    z = np.load(path, allow_pickle=False)
    g = Generator(int(z["config"]), **json.loads(str(z["overrides"])))
    assert g.digest == int(z["digest"]), "generator changed: regenerate with tests/golden/make_mvd.py"
    f = MotionField(g.width_mbs, g.height_mbs)
    n4 = n8 = rem = 0
    slot_pic = {}
    for i in range(g.num_pictures):
        desc, mbs, _ = g.picture(i)
        col = slot_pic[desc.ref_slot[1][0]] if desc.slice_type == _lib.SLICE_B else None
        got = f.derive_picture(mbs, z["mvd"][i], z["imode"][i], bool(desc.constrained_intra_pred_flag), col=col)
        slot_pic[desc.dst_slot] = got
        for a in np.nonzero(mbs["mb_type"] == _lib.MB_I4x4)[0]:
            assert (got["intra_pred_modes"][a] == mbs["intra_pred_modes"][a]).all(), (i, int(a))
            n4 += 1
        for a in np.nonzero(mbs["mb_type"] == _lib.MB_I8x8)[0]:
            assert (got["intra_pred_modes"][a, :4] == mbs["intra_pred_modes"][a, :4]).all(), (i, int(a))
            n8 += 1
        rem += int((z["imode"][i][..., 0] == 0).sum())
    assert n4 > 0 and rem > 0
    if int(z["config"]) == 2:
        assert n8 > 0


def test_isolated_mb_predicts_zero():
    """With no decoded neighbour every prediction is 0 (8.4.1.3.2): mv = mvd."""
    mbs = np.zeros(1, dtype=_lib.MB_DTYPE)
    mbs["mb_type"] = 4
    mbs["part_shape"] = 3
    mbs["sub_part_shape"] = [0, 1, 2, 3]
    mbs["ref_idx"][0, 0] = [0, 1, 0, 2]
    mbs["ref_idx"][0, 1] = -1
    mvd = np.zeros((1, 2, 16, 2), np.int16)
    mvd[0, 0, 0] = (5, -3)  # quadrant 0, 8x8
    f = MotionField(1, 1)
    got = f.derive_picture(mbs, mvd)
    assert tuple(got["mv"][0, 0, 0]) == (5, -3) and tuple(got["mv"][0, 0, 5]) == (5, -3)
    # quadrant 1 (8x4, ref 1): A is quadrant 0 (ref 0), B, C and D lie outside the
    # picture, so B and C take A's vector and reference (8.4.1.3.1) and the median is
    # A's vector although its reference differs
    assert tuple(got["mv"][0, 0, 2]) == (5, -3) and tuple(got["mv"][0, 0, 3]) == (5, -3)


def test_rejects_what_it_does_not_derive():
    mbs = np.zeros(1, dtype=_lib.MB_DTYPE)
    mbs["mb_type"] = 4
    f = MotionField(1, 1)
    with pytest.raises(_lib.H264GError):  # P MB without mvd
        _lib.check(_lib.lib().h264g_mvfield_derive(f._h, 0, mbs.ctypes.data, None))
    with pytest.raises(_lib.H264GError):  # MB address outside the picture
        _lib.check(_lib.lib().h264g_mvfield_derive(f._h, 1, mbs.ctypes.data, np.zeros(64, np.int16).ctypes.data))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cif_ip", "high_small", "high_direct_pcm"])
def test_pictures_from_derived_mvs_match_reference(golden, name):
    """A parser's path: records whose motion vectors come from the mvds through the host
    derivation (not from the generator), reconstructed on the GPU, give the reference's
    pictures."""
    from h264_video_decoder_demo_amd import Engine
    from helpers import plane_hashes
    z = np.load(os.path.join(GOLDEN, f"mvd_{name}.npz"), allow_pickle=False)
    g = Generator(int(z["config"]), **json.loads(str(z["overrides"])))
    assert g.digest == int(golden[name]["digest"], 16)  # the golden pictures are of this syntax
    f = MotionField(g.width_mbs, g.height_mbs)
    slot_pic = {}
    with Engine(device=0) as eng:
        sid = eng.open_stream(g.width_mbs, g.height_mbs, g.dpb_slots)
        for i in range(g.num_pictures):
            desc, mbs, levels = g.picture(i)
            col = slot_pic[desc.ref_slot[1][0]] if desc.slice_type == _lib.SLICE_B else None
            derived = f.derive_picture(mbs, z["mvd"][i], z["imode"][i], col=col)
            slot_pic[desc.dst_slot] = derived
            p = eng.begin(sid, desc, derived, levels)
            eng.submit(p)
            eng.sync()
            assert plane_hashes(*eng.readback(sid, desc.dst_slot)) == golden[name]["pic"][i], (name, i)
            eng.release(p)
        eng.close_stream(sid)


def _mvd_stream(tmp_path, name):
    from h264_video_decoder_demo_amd.syntax_io import SyntaxFile, save_stream
    z = np.load(os.path.join(GOLDEN, f"mvd_{name}.npz"), allow_pickle=False)
    g = Generator(int(z["config"]), **json.loads(str(z["overrides"])))
    path = str(tmp_path / f"{name}.h264g.npz")
    save_stream(g, path, mvd=list(z["mvd"]))
    return g, SyntaxFile(path)


def test_syntax_file_with_mvd_derives_the_vectors(tmp_path):
    """A stream stored with mvds and no final vectors reads back with the generator's
    vectors (CPU)."""
    g, f = _mvd_stream(tmp_path, "high_small_seed7")
    for i in range(g.num_pictures):
        _, want, _ = g.picture(i)
        _, got, _ = f.picture(i)
        used = np.repeat(np.repeat(want["ref_idx"].reshape(-1, 2, 2, 2), 2, axis=2), 2, axis=3).reshape(-1, 2, 16) >= 0
        assert (got["mv"][used] == want["mv"][used]).all(), i


@pytest.mark.gpu
def test_decoder_on_mvd_stream_matches_reference(golden, tmp_path):
    """The reference's decoder interface fed a stream with mvds (a parser that stops at
    mvd) outputs the reference's pictures."""
    from h264_video_decoder_demo_amd import H264VideoDecoder
    from helpers import plane_hashes
    g, _ = _mvd_stream(tmp_path, "cif_ip")
    got = []

    def on_frame(pic, user, err):
        if pic is not None:
            got.append(plane_hashes(pic.m_pic_buff_luma, pic.m_pic_buff_cb, pic.m_pic_buff_cr))
        return 0

    dec = H264VideoDecoder(device=0)
    dec.set_output_frame_callback_functuin(on_frame, None)
    assert dec.open(str(tmp_path / "cif_ip.h264g.npz")) == 0
    # cif_ip is I then P with no reordering: output (POC) order is decoding order
    assert got == golden["cif_ip"]["pic"]
