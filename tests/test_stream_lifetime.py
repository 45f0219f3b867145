"""A closed stream's pictures and batches are refused before any launch.

h264g_stream_close returns the stream's DPB, progress words and strips to the context,
and the stream id goes to the next stream opened.  A picture staged before the close
(or a batch built from it) still points at that memory, so submitting it, building a
batch from it or running such a batch must fail with H264G_E_STATE and launch nothing.
"""
import pytest

from h264_video_decoder_demo_amd import Engine, Generator
from h264_video_decoder_demo_amd._lib import H264G_E_STATE, H264GError

from helpers import oracle_decode, plane_hashes

pytestmark = pytest.mark.gpu


def _refused(fn):
    with pytest.raises(H264GError) as e:
        fn()
    assert e.value.code == H264G_E_STATE, e.value


def test_closed_stream_pictures_are_refused():
    g = Generator(1, width_mbs=5, height_mbs=4, n_frames=2)
    want = [plane_hashes(*out[2]) for _, _, out in oracle_decode(g)]
    with Engine(device=0) as eng:
        sid = eng.open_stream(g.width_mbs, g.height_mbs, g.dpb_slots)
        d0, m0, l0 = g.picture(0)
        stale = eng.begin(sid, d0, m0, l0)
        batch = eng.batch_create([stale])
        eng.close_stream(sid)
        _refused(lambda: eng.submit(stale))
        _refused(lambda: eng.batch_run(batch))
        _refused(lambda: eng.batch_submit([stale]))
        _refused(lambda: eng.batch_create([stale]))
        # the id now names a new stream: the old picture and batch stay refused
        sid2 = eng.open_stream(g.width_mbs, g.height_mbs, g.dpb_slots)
        assert sid2 == sid
        _refused(lambda: eng.submit(stale))
        _refused(lambda: eng.batch_run(batch))
        eng.sync()  # nothing was launched for them, so nothing can have failed
        eng.batch_destroy(batch)
        eng.release(stale)
        # the new stream decodes normally, from a clean DPB
        for i in range(g.num_pictures):
            d, m, lv = g.picture(i)
            p = eng.begin(sid2, d, m, lv)
            eng.submit(p)
            eng.sync()
            assert plane_hashes(*eng.readback(sid2, d.dst_slot)) == want[i]
            eng.release(p)
        eng.close_stream(sid2)
        assert sid2 not in eng.streams


def test_field_pictures_are_refused():
    """PAFF field pictures have no reference output (the reference crashes on P / B
    fields and misplaces I-field MBs), so the ABI refuses them before any launch."""
    from h264_video_decoder_demo_amd._lib import H264G_E_ARG
    g = Generator(1, width_mbs=5, height_mbs=4, n_frames=1)
    with Engine(device=0) as eng:
        sid = eng.open_stream(g.width_mbs, g.height_mbs, g.dpb_slots)
        d, m, lv = g.picture(0)
        for structure in (1, 2):  # H264G_TOP_FIELD, H264G_BOTTOM_FIELD
            d.structure = structure
            with pytest.raises(H264GError) as e:
                eng.begin(sid, d, m, lv)
            assert e.value.code == H264G_E_ARG, e.value
        eng.close_stream(sid)
