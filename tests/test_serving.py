"""The serving loop: every frame decoded in the mix (continuous batching, DPB budget
admission) is byte-identical to the same stream decoded alone."""
import numpy as np
import pytest

from h264_video_decoder_demo_amd import Engine, Generator
from h264_video_decoder_demo_amd.serve import ServeLoop, StreamSpec, arrival_times, make_specs

from helpers import plane_hashes


def test_specs_follow_the_mix_and_rate():
    specs = make_specs(400, total_fps=400.0, duration_s=4.0)
    sizes = [(s.width_mbs, s.height_mbs) for s in specs]
    share_1080 = sizes.count((120, 68)) / len(sizes)
    assert 0.22 < share_1080 < 0.38
    assert all(s.rate_fps == pytest.approx(1.0) for s in specs)
    assert specs == make_specs(400, total_fps=400.0, duration_s=4.0)  # deterministic


def test_arrivals_are_poisson_with_the_stream_rate():
    spec = StreamSpec(4, 4, 1, 20000, 0.5, 50.0)
    t = arrival_times(spec, np.random.default_rng(3))
    gaps = np.diff(t)
    assert t[0] == 0.5 and (gaps >= 0).all()
    assert gaps.mean() == pytest.approx(1 / 50.0, rel=0.05)
    assert gaps.std() == pytest.approx(1 / 50.0, rel=0.08)  # exponential: std = mean


def _specs():
    sizes = [(22, 18), (12, 9), (30, 20), (7, 5), (16, 12), (22, 18)]
    return [StreamSpec(w, h, 0x77 + i, 6, 0.002 * i, 400.0) for i, (w, h) in enumerate(sizes)]


def _alone(specs):
    out = {}
    with Engine(device=0) as eng:
        for i, sp in enumerate(specs):
            g = Generator(2, width_mbs=sp.width_mbs, height_mbs=sp.height_mbs, n_frames=sp.n_frames, seed=sp.seed,
                          gop=32)
            sid = eng.open_stream(g.width_mbs, g.height_mbs, g.dpb_slots)
            for k in range(g.num_pictures):
                d, m, lv = g.picture(k)
                p = eng.begin(sid, d, m, lv)
                eng.submit(p)
                eng.sync()
                out[(i, k)] = plane_hashes(*eng.readback(sid, d.dst_slot))
                eng.release(p)
            eng.close_stream(sid)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("budget_streams", [0, 2])
def test_serving_mix_equals_alone(budget_streams):
    specs = _specs()
    want = _alone(specs)
    got = {}

    def on_frame(i, k, eng, sid, desc):
        got[(i, k)] = plane_hashes(*eng.readback(sid, desc.dst_slot))

    budget = 0
    if budget_streams:  # room for about two of the largest DPBs: later streams must wait
        # slots, per-MB and per-row progress words, strips of 3 deblocking bands and 3 intra
        # bands of 8 rows
        budget = 2 * (22 * 18 * 384 * 5 + 22 * 18 * 4 + 18 * 8 + 3 * 22 * 256 + 3 * 22 * 64) + 1024
    # streams of 396 and 600 MBs run on the context's second HIP stream
    with Engine(device=0, dpb_budget_bytes=budget, lane_split_mbs=300) as eng:
        res = ServeLoop(eng, specs, window_ms=1.0, on_frame=on_frame).run()
    assert res.frames == sum(s.n_frames for s in specs)
    assert got == want
    assert len(res.latencies_ms) == res.frames
    if budget_streams:
        assert res.admitted_late > 0
    else:
        assert res.max_batch > 1  # frames of several streams shared launches
