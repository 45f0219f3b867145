"""HIP engine against the reference's golden hashes and the C oracle (MI355X only).

Every picture of every golden case is reconstructed through the C ABI
(libh264g.so, gfx950 kernels) and must be byte-identical to the reference
decoder's output.  On a mismatch the picture is decoded again with deblocking
off and compared with the reference's pre-deblocking hash, and the oracle's
planes locate the first differing macroblock.
"""
import ctypes

import numpy as np
import pytest

from h264_video_decoder_demo_amd import Generator

from helpers import first_diff, make_generator, oracle_decode, plane_hashes

pytestmark = pytest.mark.gpu

CASES = ["cif_ip", "cif_ip_seed2", "cif_ip_seed3", "cif_pcm_cqp", "high_small", "high_small_seed7",
         "high_implicit", "high_default_nodbk", "narrow", "flat", "high_1080p", "high_720p", "high_2160p",
         "mbaff_small", "mbaff_t8_pcm", "mbaff_tall", "mbaff_1080p", "high_small_cip", "high_odd_cip",
         "high_direct", "high_direct_odd", "high_direct_pcm"]


def _diagnose(engine, sid, name, gen, i, desc, mbs, levels, got):
    """Message naming the stage and location of a mismatch of picture i."""
    msg = [f"{name}: picture {i} (slice_type {desc.slice_type}) differs from the reference"]
    ref = None
    for j, _, out in oracle_decode(gen, (1, 2)):
        if j == i:
            ref = out
            break
    for k, pl in enumerate(("Y", "Cb", "Cr")):
        d = first_diff(got[k], ref[2][k], 16 if k == 0 else 8)
        if d:
            msg.append(f"  final {pl}: {d}")
    d2 = ctypes.pointer(type(desc).from_buffer_copy(desc)).contents
    d2.disable_deblocking_filter_idc = 1
    pic = engine.begin(sid, d2, mbs, levels)
    engine.submit(pic)
    engine.sync()
    pre = engine.readback(sid, desc.dst_slot)
    engine.release(pic)
    for k, pl in enumerate(("Y", "Cb", "Cr")):
        d = first_diff(pre[k], ref[1][k], 16 if k == 0 else 8)
        if d:
            msg.append(f"  before deblocking {pl}: {d}")
    return "\n".join(msg)


@pytest.mark.parametrize("name", CASES)
def test_plain_path_matches_reference(any_engine, golden, name):
    engine = any_engine
    case = golden[name]
    gen = make_generator(case)
    sid = engine.open_stream_for(gen)
    try:
        for i in range(gen.num_pictures):
            desc, mbs, levels = gen.picture(i)
            pic = engine.begin(sid, desc, mbs, levels)
            engine.submit(pic)
            engine.sync()
            got = engine.readback(sid, desc.dst_slot)
            engine.release(pic)
            if plane_hashes(*got) != case["pic"][i]:
                pytest.fail(_diagnose(engine, sid, name, gen, i, desc, mbs, levels, got))
    finally:
        engine.close_stream(sid)


def test_batched_path_matches_reference(any_engine, golden):
    """Pictures of several streams of different sizes reconstructed in one batch per
    picture index give the same bytes as the reference."""
    engine = any_engine
    names = ["cif_ip_seed2", "high_small", "high_implicit", "high_default_nodbk", "narrow", "flat", "cif_pcm_cqp",
             "high_720p", "high_2160p", "mbaff_small", "mbaff_t8_pcm", "high_odd_cip",
             "high_direct_odd"]
    gens = {n: make_generator(golden[n]) for n in names}
    sids = {n: engine.open_stream_for(g) for n, g in gens.items()}
    try:
        npic = max(g.num_pictures for g in gens.values())
        for i in range(npic):
            staged = []
            for n, g in gens.items():
                if i < g.num_pictures:
                    desc, mbs, levels = g.picture(i)
                    staged.append((n, desc, engine.begin(sids[n], desc, mbs, levels)))
            engine.batch_submit([p for _, _, p in staged])
            engine.sync()
            for n, desc, p in staged:
                got = engine.readback(sids[n], desc.dst_slot)
                assert plane_hashes(*got) == golden[n]["pic"][i], f"batched {n} picture {i}"
                engine.release(p)
    finally:
        for s in sids.values():
            engine.close_stream(s)


def test_batch_rerun_is_idempotent(engine, golden):
    """A prepared batch run twice over the same inputs gives the same output."""
    case = golden["high_small"]
    gen = make_generator(case)
    sid = engine.open_stream_for(gen)
    try:
        for i in range(2):  # I and P picture
            desc, mbs, levels = gen.picture(i)
            pic = engine.begin(sid, desc, mbs, levels)
            b = engine.batch_create([pic])
            engine.batch_run(b)
            engine.sync()
            first = engine.readback(sid, desc.dst_slot)
            assert plane_hashes(*first) == case["pic"][i]
            if i == 1:  # the P picture only reads other slots, so a re-run must give the same bytes
                engine.batch_run(b)
                engine.batch_run(b)
                engine.sync()
                again = engine.readback(sid, desc.dst_slot)
                for a, c in zip(first, again):
                    np.testing.assert_array_equal(a, c)
            engine.batch_destroy(b)
            engine.release(pic)
    finally:
        engine.close_stream(sid)


def test_graph_replays_match_eager_runs(engine, golden):
    """A batch's captured stage graph, replayed several times with eager per-picture
    launches and synchronisations in between, reproduces the reference every time."""
    names = ["high_small", "cif_ip_seed3"]
    gens = {n: make_generator(golden[n]) for n in names}
    sids = {n: engine.open_stream_for(g) for n, g in gens.items()}
    try:
        for i in range(2):  # I then P picture of both streams
            staged = [(n, g.picture(i)) for n, g in gens.items()]
            pics = [(n, d, engine.begin(sids[n], d, m, lv)) for n, (d, m, lv) in staged]
            b = engine.batch_create([p for _, _, p in pics])
            for rep in range(4):
                if rep % 2:
                    for _, _, p in pics:  # eager launches of the same pictures
                        engine.submit(p)
                else:
                    engine.batch_run(b)  # graph replay (captured on the first run)
                engine.sync()
                for n, d, _ in pics:
                    got = engine.readback(sids[n], d.dst_slot)
                    assert plane_hashes(*got) == golden[n]["pic"][i], f"{n} picture {i} run {rep}"
            engine.batch_destroy(b)
            for _, _, p in pics:
                engine.release(p)
    finally:
        for s in sids.values():
            engine.close_stream(s)


@pytest.mark.parametrize("name", ["cif_ip", "high_small", "flat", "mbaff_small"])
def test_bgr24_matches_reference(engine, golden, name):
    """Output conversion on the device against the reference's convertYuv420pToBgr24."""
    import hashlib
    case = golden[name]
    gen = make_generator(case)
    sid = engine.open_stream_for(gen)
    try:
        for i in range(gen.num_pictures):
            desc, mbs, levels = gen.picture(i)
            pic = engine.begin(sid, desc, mbs, levels)
            engine.submit(pic)
            engine.sync()
            bgr = engine.to_bgr24(sid, desc.dst_slot)
            assert hashlib.sha256(bgr.tobytes()).hexdigest() == case["bgr"][i], f"{name} picture {i}"
            flipped = engine.to_bgr24(sid, desc.dst_slot, flip=True)
            np.testing.assert_array_equal(flipped[::-1], bgr)
            engine.release(pic)
    finally:
        engine.close_stream(sid)


@pytest.mark.parametrize("width_mbs", [240, 330, 340])
def test_wide_pictures_match_oracle(any_engine, width_mbs):
    """Pictures as wide as 2160p and wider (240 to 340 MBs, a few rows): every schedule
    matches the oracle."""
    engine = any_engine
    gen = Generator(2, width_mbs=width_mbs, height_mbs=5, n_frames=3, seed=11 + width_mbs)
    want = {i: plane_hashes(*out[2]) for i, _, out in oracle_decode(gen)}
    sid = engine.open_stream_for(gen)
    try:
        for i in range(gen.num_pictures):
            desc, mbs, levels = gen.picture(i)
            pic = engine.begin(sid, desc, mbs, levels)
            engine.submit(pic)
            engine.sync()
            got = engine.readback(sid, desc.dst_slot)
            engine.release(pic)
            assert plane_hashes(*got) == want[i], f"width {width_mbs} picture {i}"
    finally:
        engine.close_stream(sid)


def test_overlapped_frames_match_oracle(any_engine):
    """Frame indices of several streams batched and run back to back with no wait between
    them.  The streams have spare DPB slots, so a picture does not overwrite the slot of
    the picture before it and a B picture's deblocking overlaps the next picture's stages
    (the engine orders a picture after a deblocking launch only when it reads or writes a
    slot that launch writes).  At the end every slot of every stream equals the oracle's
    DPB."""
    from oracle.refcpu import Dpb
    engine = any_engine
    sizes = [(22, 18, 11), (40, 23, 12), (30, 20, 13)]  # the 920-MB stream runs on the second lane
    gens = [Generator(2, width_mbs=w, height_mbs=h, n_frames=9, seed=0x0E4A0000 + s, spare_slots=2)
            for w, h, s in sizes]
    gens.append(Generator(3, width_mbs=30, height_mbs=22, n_frames=9, seed=0x0E4A0014, spare_slots=2))  # MBAFF
    sids = [engine.open_stream_for(g) for g in gens]
    try:
        frames = [[g.picture(i) for i in range(g.num_pictures)] for g in gens]
        pics = [[engine.begin(sid, *f) for f in fr] for sid, fr in zip(sids, frames)]
        batches = [engine.batch_create([pics[k][i] for k in range(len(gens)) if i < len(pics[k])])
                   for i in range(max(len(p) for p in pics))]
        for b in batches:
            engine.batch_run(b)
        engine.sync()
        for k, g in enumerate(gens):
            dpb = Dpb(g.width_mbs, g.height_mbs, g.dpb_slots)
            for desc, mbs, levels in frames[k]:
                assert dpb.decode(desc, mbs, levels, 2) == 0
            written = {desc.dst_slot for desc, _, _ in frames[k]}
            for slot in sorted(written):
                got = engine.readback(sids[k], slot)
                for gp, wp, pl in zip(got, dpb.split(slot), ("Y", "Cb", "Cr")):
                    assert np.array_equal(gp, wp), f"stream {k} slot {slot} plane {pl}"
        for b in batches:
            engine.batch_destroy(b)
        for row in pics:
            for p in row:
                engine.release(p)
    finally:
        for sid in sids:
            engine.close_stream(sid)


def test_dependency_level_batches_match_oracle(any_engine):
    """Batches of several pictures of one stream: the frame indices of four streams
    (one MBAFF, one on the second lane) grouped into runs of mutually independent
    pictures (I0 | P3 | B1 B2 P6 | B4 B5 P9 ...), so each batch holds up to three pictures
    of every stream, each with a progress set of its own.  Run twice back to back with
    no wait (the second pass re-decodes the same DPB contents); at the end every slot of
    every stream equals the oracle's DPB."""
    from oracle.refcpu import Dpb
    from h264_video_decoder_demo_amd.engine import independent_runs
    engine = any_engine
    sizes = [(22, 18, 21), (40, 23, 22), (17, 9, 23)]  # the 920-MB stream runs on the second lane
    gens = [Generator(2, width_mbs=w, height_mbs=h, n_frames=11, seed=0x0E4A0100 + s, spare_slots=2)
            for w, h, s in sizes]
    gens.append(Generator(3, width_mbs=30, height_mbs=22, n_frames=11, seed=0x0E4A0114, spare_slots=2))  # MBAFF
    sids = [engine.open_stream_for(g) for g in gens]
    try:
        frames = [[g.picture(i) for i in range(g.num_pictures)] for g in gens]
        pics = [[engine.begin(sid, *f) for f in fr] for sid, fr in zip(sids, frames)]
        runs = independent_runs(pics)
        assert max(len(r) for r in runs) == 3, runs
        batches = [engine.batch_create([row[f] for f in run for row in pics]) for run in runs]
        for _ in range(2):
            for b in batches:
                engine.batch_run(b)
        engine.sync()
        n = min(len(r) for r in pics)
        for k, g in enumerate(gens):
            dpb = Dpb(g.width_mbs, g.height_mbs, g.dpb_slots)
            for desc, mbs, levels in frames[k][:n]:
                assert dpb.decode(desc, mbs, levels, 2) == 0
            written = {desc.dst_slot for desc, _, _ in frames[k][:n]}
            for slot in sorted(written):
                got = engine.readback(sids[k], slot)
                for gp, wp, pl in zip(got, dpb.split(slot), ("Y", "Cb", "Cr")):
                    assert np.array_equal(gp, wp), f"stream {k} slot {slot} plane {pl}"
        for b in batches:
            engine.batch_destroy(b)
        for row in pics:
            for p in row:
                engine.release(p)
    finally:
        for sid in sids:
            engine.close_stream(sid)


def test_dependent_pictures_of_one_stream_are_refused(engine):
    """A batch refuses two pictures of one stream that depend on each other (one reads or
    writes the slot the other writes), before anything is launched."""
    from h264_video_decoder_demo_amd._lib import H264GError
    gen = Generator(2, width_mbs=6, height_mbs=4, n_frames=9, seed=0x0E4A0200, spare_slots=2)
    sid = engine.open_stream_for(gen)
    try:
        pics = [engine.begin(sid, *gen.picture(i)) for i in range(gen.num_pictures)]
        for pair in ([0, 1], [1, 2], [2, 3, 4, 5]):  # P3 reads I0; B1 reads P3; B4 reads P6
            with pytest.raises(H264GError, match="depend on each other"):
                engine.batch_create([pics[i] for i in pair])
        b = engine.batch_create([pics[i] for i in (2, 3, 4)])  # B1 B2 P6
        engine.batch_destroy(b)
        for p in pics:
            engine.release(p)
    finally:
        engine.close_stream(sid)


def _decode_all(engine, gen, pictures):
    """Every (desc, mbs, levels) of pictures through the engine, compared with the oracle
    decoding the same records on its own DPB."""
    from oracle.refcpu import Dpb
    dpb = Dpb(gen.width_mbs, gen.height_mbs, gen.dpb_slots)
    sid = engine.open_stream_for(gen)
    try:
        for i, (desc, mbs, levels) in enumerate(pictures):
            assert dpb.decode(desc, mbs, levels, 2) == 0
            pic = engine.begin(sid, desc, mbs, levels)
            engine.submit(pic)
            engine.sync()
            got = engine.readback(sid, desc.dst_slot)
            engine.release(pic)
            for gp, wp, pl in zip(got, dpb.split(desc.dst_slot), ("Y", "Cb", "Cr")):
                d = first_diff(gp, wp, 16 if pl == "Y" else 8)
                assert d is None, f"picture {i} {pl}: {d}"
    finally:
        engine.close_stream(sid)


@pytest.mark.parametrize("slices,idc", [(7, 2), (23, 2), (3, 1), (7, 0), (23, 0)])
def test_multi_slice_pictures_match_oracle(any_engine, slices, idc):
    """Pictures of several slices (runs of MBs that end mid-row): intra prediction sees no
    neighbour of another slice (6.4.8); with disable_deblocking_filter_idc 2 MB edges
    between slices are not filtered, with 0 they are, with the current MB in the
    neighbour's place for bS and qP (quirk Q7); I, P and B pictures with a third of the
    inter pictures' MBs intra, against the oracle."""
    gen = Generator(2, width_mbs=23, height_mbs=9, n_frames=5, seed=0x51CE0000 + slices, slices=slices,
                    disable_deblocking_filter_idc=idc, intra_in_inter_permille=330)
    pics = [gen.picture(i) for i in range(gen.num_pictures)]
    for _, mbs, _ in pics:
        assert len(np.unique(mbs["slice_num"])) == slices
    _decode_all(any_engine, gen, pics)


def test_scaling_matrices_match_oracle(any_engine):
    """Pictures dequantised with scaling matrices (random weights 1..255 per picture and
    list): 4x4 and 8x8 luma blocks of intra and inter MBs, Intra16x16 DC, chroma DC and AC
    (which take the MB's luma list, quirk Q13), against the oracle."""
    gen = Generator(2, width_mbs=19, height_mbs=7, n_frames=5, seed=0x5CA1E000, intra_in_inter_permille=300)
    rng = np.random.default_rng(0x5CA1E)
    pics = []
    for i in range(gen.num_pictures):
        desc, mbs, levels = gen.picture(i)
        desc.scaling_matrix = 1
        w4, w8 = rng.integers(1, 256, (6, 16)), rng.integers(1, 256, (2, 64))
        for l in range(6):
            for k in range(16):
                desc.weight_scale4x4[l][k] = int(w4[l, k])
        for l in range(2):
            for k in range(64):
                desc.weight_scale8x8[l][k] = int(w8[l, k])
        pics.append((desc, mbs, levels))
    _decode_all(any_engine, gen, pics)


def _lose_tail(desc, mbs, levels, first):
    """The picture with MBs first.. replaced by what a damaged slice leaves: zero I_PCM
    records of no slice."""
    from h264_video_decoder_demo_amd import _lib
    mbs = mbs.copy()
    tail = mbs[first:]
    tail[...] = 0
    tail["mb_type"] = _lib.MB_IPCM
    tail["flags"] = _lib.MBF_LOST
    tail["slice_num"] = _lib.SLICE_NONE
    tail["ref_idx"] = -1
    tail["level_offset"] = len(levels) + 384 * np.arange(len(tail))
    return desc, mbs, np.concatenate([levels, np.zeros(384 * len(tail), np.int16)])


def test_mbs_of_no_slice_stop_deblocking(any_engine):
    """MBs that no slice decoded (zero I_PCM records, slice H264G_SLICE_NONE) from a
    mid-row MB on: deblocking stops at the first of them, so neither its edges nor any
    later MB's are filtered, in I, P and B pictures (single-slice pictures with
    disable_deblocking_filter_idc 0 and 2-slice ones with 2), against the oracle."""
    for idc, slices in ((0, 1), (2, 2)):
        gen = Generator(2, width_mbs=21, height_mbs=8, n_frames=4, seed=0x51CE0100 + idc, slices=slices,
                        disable_deblocking_filter_idc=idc)
        n = gen.width_mbs * gen.height_mbs
        pics = [_lose_tail(*gen.picture(i), first=n - 1 - (37 * i + 50) % (n // 2)) for i in range(gen.num_pictures)]
        _decode_all(any_engine, gen, pics)


def test_slice_layouts_the_engine_does_not_take_are_refused(engine):
    """Several slices, or MBs of no slice, in an MBAFF frame are refused before anything
    runs."""
    from h264_video_decoder_demo_amd import _lib
    gen = Generator(3, width_mbs=6, height_mbs=4, n_frames=1, seed=0x51CE0201)
    desc, mbs, levels = gen.picture(0)
    sid = engine.open_stream_for(gen)
    try:
        two = mbs.copy()
        two["slice_num"][12:] = 1
        with pytest.raises(_lib.H264GError, match="several slices in an MBAFF"):
            engine.begin(sid, desc, two, levels)
        with pytest.raises(_lib.H264GError, match="no slice in an MBAFF"):
            engine.begin(sid, *_lose_tail(desc, mbs, levels, first=20))
    finally:
        engine.close_stream(sid)


@pytest.mark.parametrize("shape", [(240, 6), (120, 12), (3, 40)])
def test_mbaff_pictures_match_oracle(any_engine, shape):
    """MBAFF frames of other shapes (as wide as 2160p, narrow and tall) against the
    oracle, which reproduces the reference's MBAFF golden cases."""
    engine = any_engine
    w, h = shape
    gen = Generator(3, width_mbs=w, height_mbs=h, n_frames=5, seed=0x3AFF0000 + w)
    want = {i: plane_hashes(*out[2]) for i, _, out in oracle_decode(gen)}
    sid = engine.open_stream_for(gen)
    try:
        for i in range(gen.num_pictures):
            desc, mbs, levels = gen.picture(i)
            pic = engine.begin(sid, desc, mbs, levels)
            engine.submit(pic)
            engine.sync()
            got = engine.readback(sid, desc.dst_slot)
            engine.release(pic)
            assert plane_hashes(*got) == want[i], f"MBAFF {w}x{h} picture {i}"
    finally:
        engine.close_stream(sid)
