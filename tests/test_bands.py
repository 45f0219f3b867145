"""Band heights chosen at run time (h264g_set_bands) and the band-height autotuner.

Every band height must give the reference's bytes; the autotuner only picks among
them.  GPU tests call the library through its C ABI; the search logic of the
autotuner is also checked on the CPU with a stand-in engine."""
import numpy as np
import pytest

from h264_video_decoder_demo_amd import Engine, _lib
from h264_video_decoder_demo_amd import autotune as at

from helpers import make_generator, plane_hashes

NAMES = ["cif_ip_seed2", "high_small", "narrow", "flat"]


def _run_batches(engine, gens, sids, golden):
    npic = max(g.num_pictures for g in gens.values())
    for i in range(npic):
        staged = []
        for n, g in gens.items():
            if i < g.num_pictures:
                desc, mbs, levels = g.picture(i)
                staged.append((n, desc, engine.begin(sids[n], desc, mbs, levels)))
        b = engine.batch_create([p for _, _, p in staged])
        engine.batch_run(b)
        engine.sync()
        for n, desc, p in staged:
            assert plane_hashes(*engine.readback(sids[n], desc.dst_slot)) == golden[n]["pic"][i], (n, i)
            engine.release(p)
        engine.batch_destroy(b)


@pytest.mark.gpu
def test_set_bands_changes_heights_and_keeps_bytes(golden):
    with Engine(device=0, deblock_waves=4, intra_waves=3, lane_split_mbs=500) as eng:
        gens = {n: make_generator(golden[n]) for n in NAMES}
        sids = {n: eng.open_stream(g.width_mbs, g.height_mbs, g.dpb_slots) for n, g in gens.items()}
        # strips were sized for 4 / 3 rows: lower heights are refused
        with pytest.raises(_lib.H264GError) as e:
            eng.set_bands(2, 3)
        assert e.value.code == _lib.H264G_E_STATE
        for d, i in ((4, 3), (7, 8), (16, 5), (24, 8)):
            eng.set_bands(d, i)
            _run_batches(eng, gens, sids, golden)
        for s in sids.values():
            eng.close_stream(s)


@pytest.mark.gpu
def test_batch_of_other_heights_is_refused(golden):
    with Engine(device=0) as eng:
        g = make_generator(golden["cif_ip"])
        sid = eng.open_stream(g.width_mbs, g.height_mbs, g.dpb_slots)
        desc, mbs, levels = g.picture(0)
        p = eng.begin(sid, desc, mbs, levels)
        b = eng.batch_create([p])
        eng.set_bands(12, 8)
        with pytest.raises(_lib.H264GError):
            eng.batch_run(b)
        eng.batch_destroy(b)
        b = eng.batch_create([p])
        eng.batch_run(b)
        eng.sync()
        assert plane_hashes(*eng.readback(sid, desc.dst_slot)) == golden["cif_ip"]["pic"][0]
        eng.batch_destroy(b)
        eng.release(p)
        eng.close_stream(sid)


@pytest.mark.gpu
def test_autotune_picks_a_candidate(golden, tmp_path, monkeypatch):
    monkeypatch.setenv("XDG_CACHE_HOME", str(tmp_path))
    gens = [make_generator(golden[n]) for n in NAMES]
    r = at.autotune(gens, n_frames=3, reps=1, deblock_rows=(4, 8), intra_rows=(4, 8))
    looked_up = at.autotune(gens, n_frames=3, reps=1, deblock_rows=(4, 6), intra_rows=(4, 8), measure=False)
    assert not looked_up["cached"] and (looked_up["deblock_waves"], looked_up["intra_waves"]) == (8, 8)
    assert r["deblock_waves"] in (4, 8) and r["intra_waves"] in (4, 8) and not r["cached"]
    assert set(r["ms"]) == {"4/8", "8/8", f"{r['deblock_waves']}/4"}
    again = at.autotune(gens, n_frames=3, reps=1, deblock_rows=(4, 8), intra_rows=(4, 8))
    assert again["cached"] and (again["deblock_waves"], again["intra_waves"]) == (r["deblock_waves"], r["intra_waves"])


class _FakeEngine:
    """Stands in for Engine on the CPU: a batch run 'takes' a time set by the heights."""
    cost = {}

    def __init__(self, device=0, **kw):
        self.bands = (kw["deblock_waves"], kw["intra_waves"])
        self.t = 0.0

    def open_stream(self, w, h, slots):
        return 0

    def begin_raw(self, sid, i):
        from h264_video_decoder_demo_amd._lib import PicDesc
        d = PicDesc()  # picture i writes slot i and reads slot i - 1: one batch per picture
        d.dst_slot = i
        for l in range(2):
            for r in range(32):
                d.ref_slot[l][r] = -1
        d.ref_slot[0][0] = i - 1
        return d

    def set_bands(self, d, i):
        assert d >= self.bands[0] and i >= self.bands[1]  # never below the strips' sizing
        self.cur = (d, i)

    def batch_create(self, pics):
        return self.cur

    def batch_run(self, b):
        assert b == self.cur
        _FakeEngine.clock[0] += self.cost[b]

    def batch_destroy(self, b):
        pass

    def sync(self):
        pass

    def close(self):
        pass


class _Gen:
    width_mbs, height_mbs, dpb_slots, num_pictures = 4, 3, 2, 4

    def picture_raw(self, i):
        return (i,)


def test_autotune_search_on_cpu(monkeypatch, tmp_path):
    """Coordinate search over a known cost table: the deblocking height first (intra
    at 8), then the intra height at the chosen deblocking height; the cache returns
    the same choice without measuring."""
    import h264_video_decoder_demo_amd.engine as engine_mod
    clock = [0.0]
    _FakeEngine.clock = clock
    _FakeEngine.cost = {(d, i): abs(d - 10) + abs(i - 6) * 0.5 + 1 for d in at.DEBLOCK_ROWS for i in at.INTRA_ROWS}
    monkeypatch.setattr(engine_mod, "Engine", _FakeEngine)
    monkeypatch.setattr(at, "_clock", lambda: clock[0])
    monkeypatch.setattr(at, "_device_id", lambda device: "gfx950|test device|256 CUs|2400000 kHz")
    monkeypatch.setenv("XDG_CACHE_HOME", str(tmp_path))
    r = at.autotune([_Gen(), _Gen()], n_frames=2, reps=2)
    assert (r["deblock_waves"], r["intra_waves"]) == (10, 6)
    assert len(r["ms"]) == len(at.DEBLOCK_ROWS) + len(at.INTRA_ROWS) - 1
    again = at.autotune([_Gen(), _Gen()], n_frames=2, reps=2)
    assert again["cached"] and (again["deblock_waves"], again["intra_waves"]) == (10, 6)


def test_autotune_cache_key_names_device_build_and_settings(monkeypatch):
    """A choice measured on another device, another library build or with other engine
    settings is not reused."""
    dev = ["gfx950|a|256 CUs|2400000 kHz"]
    lib_id = ["0123"]
    monkeypatch.setattr(at, "_device_id", lambda device: dev[0])
    monkeypatch.setattr(at, "_library_id", lambda: lib_id[0])
    gens = [_Gen(), _Gen()]
    k0 = at._key(gens, 2, at.DEBLOCK_ROWS, at.INTRA_ROWS, 0, {"graphs": True})
    assert k0 == at._key(gens, 2, at.DEBLOCK_ROWS, at.INTRA_ROWS, 0, {"graphs": True, "deblock_waves": 4})
    assert k0 != at._key(gens, 2, at.DEBLOCK_ROWS, at.INTRA_ROWS, 0, {"graphs": False})
    dev[0] = "gfx950|a|240 CUs|2400000 kHz"
    assert k0 != at._key(gens, 2, at.DEBLOCK_ROWS, at.INTRA_ROWS, 0, {"graphs": True})
    dev[0] = "gfx950|a|256 CUs|2400000 kHz"
    lib_id[0] = "4567"
    assert k0 != at._key(gens, 2, at.DEBLOCK_ROWS, at.INTRA_ROWS, 0, {"graphs": True})
