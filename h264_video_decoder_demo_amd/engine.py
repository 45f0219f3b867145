"""Host-side handles over the libh264g C ABI: engine context, streams, staged
pictures, batches, and the deterministic syntax generator.

Everything here is plumbing around ``include/h264g.h``; the reconstruction
itself runs in the gfx950 kernels of the library.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_float, c_uint32, c_uint64, c_void_p
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import MB_DTYPE, MB_I4x4, MB_I8x8, Config, GenParams, PicDesc, StreamDesc, check, lib


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data) if a.size else None


# ------------------------------------------------------------------ generator
class Generator:
    """Deterministic synthetic syntax for one stream (pictures in decoding order)."""

    def __init__(self, config: int = 1, seed=None, **overrides):
        self.params = GenParams()
        check(lib().h264g_gen_default_params(config, byref(self.params)))
        if seed is not None:
            self.params.seed = seed
        for k, v in overrides.items():
            if not hasattr(self.params, k):
                raise AttributeError(f"unknown generator parameter {k}")
            setattr(self.params, k, v)
        h = c_void_p()
        check(lib().h264g_gen_create(byref(self.params), byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().h264g_gen_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def width_mbs(self) -> int:
        return self.params.width_mbs

    @property
    def height_mbs(self) -> int:
        return self.params.height_mbs

    @property
    def mbaff(self) -> bool:
        """The pictures are MBAFF frames."""
        return bool(self.params.mbaff)

    @property
    def num_pictures(self) -> int:
        return check(lib().h264g_gen_num_pictures(self._h))

    @property
    def dpb_slots(self) -> int:
        return check(lib().h264g_gen_dpb_slots(self._h))

    @property
    def digest(self) -> int:
        return lib().h264g_gen_digest(self._h)

    @property
    def num_reorder_frames(self) -> int:
        """Output reorder depth of the stream (B pictures between anchors)."""
        return self.params.b_frames

    def is_idr(self, idx: int) -> bool:
        """The generator codes every I picture as an IDR picture."""
        desc = PicDesc()
        mbs_p, nmb = c_void_p(), c_uint32()
        lv_p, nlv = c_void_p(), c_uint64()
        check(lib().h264g_gen_picture(self._h, idx, byref(desc), byref(mbs_p), byref(nmb), byref(lv_p), byref(nlv)))
        return desc.slice_type == _lib.SLICE_I

    def picture(self, idx: int):
        """(PicDesc, mbs structured array, levels int16 array), copies owned by the caller."""
        desc = PicDesc()
        mbs_p, nmb = c_void_p(), c_uint32()
        lv_p, nlv = c_void_p(), c_uint64()
        check(lib().h264g_gen_picture(self._h, idx, byref(desc), byref(mbs_p), byref(nmb), byref(lv_p), byref(nlv)))
        raw = ctypes.string_at(mbs_p.value, nmb.value * MB_DTYPE.itemsize)
        mbs = np.frombuffer(raw, MB_DTYPE).copy()
        if nlv.value:
            levels = np.frombuffer(ctypes.string_at(lv_p.value, nlv.value * 2), np.int16).copy()
        else:
            levels = np.zeros(0, np.int16)
        return desc, mbs, levels

    def pictures(self):
        return [self.picture(i) for i in range(self.num_pictures)]

    def picture_raw(self, idx: int):
        """(PicDesc, mbs pointer, n_mbs, levels pointer, n_levels) into the generator's own
        storage (valid while the generator lives); no copies."""
        desc = PicDesc()
        mbs_p, nmb = c_void_p(), c_uint32()
        lv_p, nlv = c_void_p(), c_uint64()
        check(lib().h264g_gen_picture(self._h, idx, byref(desc), byref(mbs_p), byref(nmb), byref(lv_p), byref(nlv)))
        return desc, mbs_p, nmb.value, lv_p, nlv.value


class MotionField:
    """Host motion-vector derivation for one picture being parsed (h264g_mvfield_*):
    P_Skip vectors and mvp + mvd for P / B partitions of frame pictures, in the
    reference's order (H264InterPrediction.cpp:685-738, 1660-2001)."""

    def __init__(self, width_mbs: int, height_mbs: int):
        h = c_void_p()
        check(lib().h264g_mvfield_create(width_mbs, height_mbs, byref(h)))
        self._h = h
        self.width_mbs, self.height_mbs = width_mbs, height_mbs

    def close(self):
        if self._h:
            lib().h264g_mvfield_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        """Start of a picture."""
        check(lib().h264g_mvfield_reset(self._h))

    def derive_picture(self, mbs: np.ndarray, mvd: np.ndarray, imode: np.ndarray = None,
                       constrained_intra_pred: bool = False, col: np.ndarray = None,
                       col_short_term: bool = True) -> np.ndarray:
        """Records with their mv filled in: mbs (MB_DTYPE, raster order; mv ignored),
        mvd int16 [MB][list][16][2] (each partition's mvd at its top-left 4x4 block).
        imode, when given: uint8 [MB][16][2] intra-mode syntax (prev flag, rem) of the
        I_NxN MBs, whose intra_pred_modes are then derived too (8.3.1.1 / 8.3.2.1).
        col: the final records of RefPicList1[0] for the spatial-direct MBs of a B
        picture (their ref_idx and mv are then derived as well)."""
        assert mbs.dtype == MB_DTYPE and mvd.dtype == np.int16 and mvd.shape == (len(mbs), 2, 16, 2)
        assert len(mbs) == self.width_mbs * self.height_mbs
        out = mbs.copy()
        out["mv"] = 0
        mvd = np.ascontiguousarray(mvd)
        if imode is not None:
            assert imode.dtype == np.uint8 and imode.shape == (len(mbs), 16, 2)
            imode = np.ascontiguousarray(imode)
        self.reset()
        if col is not None:
            assert col.dtype == MB_DTYPE and len(col) == len(mbs)
            col = np.ascontiguousarray(col)
            check(lib().h264g_mvfield_colocated(self._h, col.ctypes.data, 1 if col_short_term else 0))
        else:
            check(lib().h264g_mvfield_colocated(self._h, None, 0))
        base = out.ctypes.data
        for a in range(len(out)):
            rec = base + a * MB_DTYPE.itemsize
            if imode is not None and out["mb_type"][a] in (MB_I4x4, MB_I8x8):
                out["intra_pred_modes"][a] = 0
                check(lib().h264g_mvfield_intra_modes(self._h, a, rec, imode[a].ctypes.data,
                                                      1 if constrained_intra_pred else 0))
            check(lib().h264g_mvfield_derive(self._h, a, rec, mvd[a].ctypes.data))
        return out


# ------------------------------------------------------------------ engine
@dataclass
class StagedPicture:
    handle: c_void_p
    stream: int
    desc: PicDesc
    width_mbs: int
    height_mbs: int


class Engine:
    """One libh264g context on one HIP device (one HIP stream inside)."""

    def __init__(self, device: int = 0, dpb_budget_bytes: int = 0, profile: bool = False, waves: int = 0,
                 max_batch_pictures: int = 0, graphs: bool = True, deblock_schedule: int = 0, deblock_waves: int = 0,
                 intra_schedule: int = 0, intra_waves: int = 0, lane_split_mbs: int = 0):
        """deblock_schedule: _lib.DEBLOCK_BANDS (default), DEBLOCK_MB_WAVEFRONT or DEBLOCK_MB_ROWS;
        intra_schedule: _lib.INTRA_BANDS (default) or INTRA_MB_WAVEFRONT.  deblock_waves /
        intra_waves: MB rows per band of the band schedules (0 = 8 / 8).  lane_split_mbs:
        streams with pictures of at least this many MBs run on the context's second HIP
        stream (0 = 16384, negative = one stream)."""
        cfg = Config(device=device, dpb_budget_bytes=dpb_budget_bytes, max_batch_pictures=max_batch_pictures,
                     profile=1 if profile else 0, waves=waves, no_graphs=0 if graphs else 1,
                     deblock_schedule=deblock_schedule, deblock_waves=deblock_waves,
                     intra_schedule=intra_schedule, intra_waves=intra_waves, lane_split_mbs=lane_split_mbs)
        h = c_void_p()
        check(lib().h264g_create(byref(cfg), byref(h)))
        self._h = h
        self.streams = {}

    def close(self):
        if self._h:
            lib().h264g_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- streams
    def open_stream(self, width_mbs: int, height_mbs: int, dpb_slots: int, mbaff: bool = False) -> int:
        """mbaff: the stream's pictures are MBAFF frames (frame_mbs_only_flag 0)."""
        d = StreamDesc(pic_width_in_mbs=width_mbs, frame_height_in_mbs=height_mbs, chroma_array_type=1, bit_depth=8,
                       frame_mbs_only=0 if mbaff else 1, mbaff=1 if mbaff else 0, dpb_slots=dpb_slots)
        sid = c_uint32()
        check(lib().h264g_stream_open(self._h, byref(d), byref(sid)))
        self.streams[sid.value] = (width_mbs, height_mbs, dpb_slots)
        return sid.value

    def open_stream_for(self, gen: "Generator") -> int:
        """A stream for the pictures of a generator (its size, DPB and frame coding)."""
        return self.open_stream(gen.width_mbs, gen.height_mbs, gen.dpb_slots, gen.mbaff)

    def close_stream(self, sid: int):
        check(lib().h264g_stream_close(self._h, sid))
        self.streams.pop(sid, None)

    def set_bands(self, deblock_rows: int = 0, intra_rows: int = 0):
        """MB rows per band of the band schedules from now on (0 keeps one); batches
        created before must be created again."""
        check(lib().h264g_set_bands(self._h, deblock_rows, intra_rows))

    # ---- pictures
    def begin(self, sid: int, desc: PicDesc, mbs: np.ndarray, levels: np.ndarray) -> StagedPicture:
        mbs = np.ascontiguousarray(mbs, MB_DTYPE)
        levels = np.ascontiguousarray(levels, np.int16)
        h = c_void_p()
        check(lib().h264g_picture_begin(self._h, sid, byref(desc), _ptr(mbs), mbs.size, _ptr(levels), levels.size,
                                        byref(h)))
        W, H, _ = self.streams[sid]
        return StagedPicture(h, sid, desc, W, H)

    def begin_raw(self, sid: int, desc: PicDesc, mbs_p, n_mbs: int, lv_p, n_levels: int) -> StagedPicture:
        """picture_begin on raw host pointers (e.g. from Generator.picture_raw)."""
        h = c_void_p()
        check(lib().h264g_picture_begin(self._h, sid, byref(desc), mbs_p, n_mbs, lv_p, n_levels, byref(h)))
        W, H, _ = self.streams[sid]
        return StagedPicture(h, sid, desc, W, H)

    def decode_static(self, sid: int, desc: PicDesc, mbs: np.ndarray, levels: np.ndarray):
        """Stages and enqueues one picture without allocating device memory
        (h264g_picture_decode_static: the device's static scratch array holds the staging)."""
        mbs = np.ascontiguousarray(mbs, MB_DTYPE)
        levels = np.ascontiguousarray(levels, np.int16)
        check(lib().h264g_picture_decode_static(self._h, sid, byref(desc), _ptr(mbs), mbs.size, _ptr(levels),
                                                levels.size))

    def submit(self, pic: StagedPicture):
        check(lib().h264g_picture_submit(self._h, pic.handle))

    def release(self, pic: StagedPicture):
        if pic.handle:
            check(lib().h264g_picture_release(self._h, pic.handle))
            pic.handle = None

    def alg_bytes(self, pic: StagedPicture):
        out = (c_uint64 * 3)()
        check(lib().h264g_picture_alg_bytes(pic.handle, out))
        return list(out)

    # ---- batches
    def batch_submit(self, pics):
        arr = (c_void_p * len(pics))(*[p.handle.value for p in pics])
        check(lib().h264g_batch_submit(self._h, arr, len(pics)))

    def batch_create(self, pics) -> c_void_p:
        arr = (c_void_p * len(pics))(*[p.handle.value for p in pics])
        b = c_void_p()
        check(lib().h264g_batch_create(self._h, arr, len(pics), byref(b)))
        return b

    def batch_run(self, b: c_void_p):
        check(lib().h264g_batch_run(self._h, b))

    def batch_destroy(self, b: c_void_p):
        check(lib().h264g_batch_destroy(self._h, b))

    # ---- completion and data
    def sync(self):
        check(lib().h264g_sync(self._h))

    def fence(self) -> c_void_p:
        """A fence after everything enqueued so far (poll with done(), free with drop_fence())."""
        f = c_void_p()
        check(lib().h264g_fence_record(self._h, byref(f)))
        return f

    def done(self, f: c_void_p) -> bool:
        return check(lib().h264g_fence_query(self._h, f)) == 1

    def wait(self, f: c_void_p):
        check(lib().h264g_fence_wait(self._h, f))

    def drop_fence(self, f: c_void_p):
        lib().h264g_fence_destroy(self._h, f)

    def set_profile(self, on: bool):
        check(lib().h264g_set_profile(self._h, 1 if on else 0))

    def stage_times(self):
        """(runs, [inter_ms, intra_ms, bs_ms, deblock_ms]) summed since the previous call."""
        out = (c_float * 4)()
        runs = check(lib().h264g_stage_times(self._h, out))
        return runs, list(out)

    def readback(self, sid: int, slot: int):
        W, H, _ = self.streams[sid]
        y = np.empty((H * 16, W * 16), np.uint8)
        cb = np.empty((H * 8, W * 8), np.uint8)
        cr = np.empty((H * 8, W * 8), np.uint8)
        check(lib().h264g_slot_readback(self._h, sid, slot, _ptr(y), _ptr(cb), _ptr(cr)))
        return y, cb, cr

    def to_bgr24(self, sid: int, slot: int, flip: bool = False) -> np.ndarray:
        """BGR24 image of a DPB slot, rows padded to 4 bytes (the reference's BMP layout)."""
        W, H, _ = self.streams[sid]
        stride = (W * 16 * 3 + 3) // 4 * 4
        out = np.zeros((H * 16, stride), np.uint8)
        check(lib().h264g_slot_to_bgr24(self._h, sid, slot, _ptr(out), stride, 1 if flip else 0))
        return out

    def upload(self, sid: int, slot: int, y: np.ndarray, cb: np.ndarray, cr: np.ndarray):
        y, cb, cr = (np.ascontiguousarray(a, np.uint8) for a in (y, cb, cr))
        check(lib().h264g_slot_upload(self._h, sid, slot, _ptr(y), _ptr(cb), _ptr(cr)))


MAX_STREAM_PICS = 4  # H264G_MAX_STREAM_PICS: pictures of one stream in one batch


def slot_use(desc: PicDesc):
    """(DPB slot the picture writes, set of slots it reads)."""
    refs = {desc.ref_slot[l][r] for l in range(2) for r in range(32) if desc.ref_slot[l][r] >= 0}
    return desc.dst_slot, refs


def independent_runs(rows, max_per_stream: int = MAX_STREAM_PICS):
    """Groups the frame indices of per-stream picture lists (StagedPicture or PicDesc, in
    decoding order) into runs of consecutive indices that can share one batch: in no
    stream does a picture of a run write a DPB slot that another picture of the run writes
    or reads (so none depends on another), and a run holds at most max_per_stream pictures
    of a stream.  With I0 P3 B1 B2 P6 B4 B5 P9 ... in decoding order and non-reference B
    pictures, the runs are {I0} {P3} {B1 B2 P6} {B4 B5 P9} ...: the batched launches of
    one stream's pictures follow its dependency levels instead of its decoding order."""
    n = min(len(r) for r in rows) if rows else 0
    uses = [[slot_use(getattr(p, "desc", p)) for p in r[:n]] for r in rows]
    runs, cur = [], []
    for f in range(n):
        ok = len(cur) < max_per_stream
        for su in uses:
            if not ok:
                break
            dst, refs = su[f]
            for g in cur:
                d2, r2 = su[g]
                if dst == d2 or d2 in refs or dst in r2:
                    ok = False
                    break
        if cur and not ok:
            runs.append(cur)
            cur = []
        cur.append(f)
    if cur:
        runs.append(cur)
    return runs


def decode_stream_plain(eng: Engine, gen: Generator, readback: bool = True):
    """Decodes a generated stream picture by picture (plain path).

    Returns the reconstructed (Y, Cb, Cr) of every picture in decoding order."""
    sid = eng.open_stream_for(gen)
    out = []
    try:
        for i in range(gen.num_pictures):
            desc, mbs, levels = gen.picture(i)
            pic = eng.begin(sid, desc, mbs, levels)
            eng.submit(pic)
            eng.sync()
            if readback:
                out.append(eng.readback(sid, desc.dst_slot))
            eng.release(pic)
    finally:
        eng.close_stream(sid)
    return out


__all__ = ["Generator", "Engine", "StagedPicture", "decode_stream_plain", "independent_runs", "slot_use",
           "MAX_STREAM_PICS", "_lib"]
