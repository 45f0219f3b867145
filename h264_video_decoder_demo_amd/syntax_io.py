"""Per-picture syntax streams on disk (.h264g.npz): what a host parser hands to the
engine, stored so that a stream can be decoded again without its generator.

Layout (numpy .npz, no pickled objects): ``meta`` int32 [width_mbs, height_mbs,
dpb_slots, num_reorder_frames, n_pictures]; ``descs`` uint8 [n, sizeof(h264g_pic_desc)];
``mbs`` h264g_mb records of all pictures; ``mb_off`` / ``lv_off`` int64 [n + 1];
``levels`` int16; ``idr`` uint8 [n]; optionally ``mvd`` int16 [MBs of all pictures][list][16][2]:
the parsed motion-vector differences, each (sub-)partition's at its top-left 4x4 block.
A stream with ``mvd`` stores its records without final motion vectors, as a parser that
stops at mvd would hand them over; reading it derives them (h264g_mvfield: P_Skip and
mvp + mvd in the reference's order).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import MB_DTYPE, PicDesc


def save_stream(source, path: str, mvd=None) -> None:
    """mvd: per picture the int16 [MB][list][16][2] motion-vector differences; when given,
    the records are stored without their motion vectors (see the module docstring)."""
    descs, mbs, lvs, idr = [], [], [], []
    for i in range(source.num_pictures):
        d, m, lv = source.picture(i)
        descs.append(np.frombuffer(bytes(d), np.uint8))
        if mvd is not None:
            m = m.copy()
            m["mv"] = 0
        mbs.append(m)
        lvs.append(lv)
        idr.append(1 if source.is_idr(i) else 0)
    mb_off = np.concatenate([[0], np.cumsum([len(m) for m in mbs])]).astype(np.int64)
    lv_off = np.concatenate([[0], np.cumsum([len(v) for v in lvs])]).astype(np.int64)
    meta = np.array([source.width_mbs, source.height_mbs, source.dpb_slots, source.num_reorder_frames,
                     source.num_pictures], np.int32)
    extra = {} if mvd is None else {"mvd": np.concatenate([np.asarray(v, np.int16) for v in mvd])}
    with open(path, "wb") as f:
        np.savez(f, meta=meta, descs=np.stack(descs), mbs=np.concatenate(mbs), mb_off=mb_off,
                 levels=np.concatenate(lvs).astype(np.int16), lv_off=lv_off, idr=np.array(idr, np.uint8), **extra)


class SyntaxFile:
    """Reads a stream written by save_stream; same interface as Generator."""

    def __init__(self, path: str):
        with np.load(path, allow_pickle=False) as z:
            self._meta = z["meta"]
            self._descs = z["descs"]
            self._mbs = z["mbs"]
            self._mb_off = z["mb_off"]
            self._levels = z["levels"]
            self._lv_off = z["lv_off"]
            self._idr = z["idr"]
            self._mvd = z["mvd"] if "mvd" in z.files else None
        if self._mbs.dtype != MB_DTYPE or self._descs.shape[1] != ctypes.sizeof(PicDesc):
            raise ValueError(f"{path}: not a syntax stream of this ABI version")
        if self._mvd is not None and self._mvd.shape != (len(self._mbs), 2, 16, 2):
            raise ValueError(f"{path}: mvd does not match the records")
        self._mvf = None

    width_mbs = property(lambda self: int(self._meta[0]))
    height_mbs = property(lambda self: int(self._meta[1]))
    dpb_slots = property(lambda self: int(self._meta[2]))
    num_reorder_frames = property(lambda self: int(self._meta[3]))
    num_pictures = property(lambda self: int(self._meta[4]))

    def is_idr(self, i: int) -> bool:
        return bool(self._idr[i])

    def picture(self, i: int):
        desc = PicDesc.from_buffer_copy(self._descs[i].tobytes())
        mbs = self._mbs[self._mb_off[i]:self._mb_off[i + 1]].copy()
        levels = self._levels[self._lv_off[i]:self._lv_off[i + 1]].copy()
        if self._mvd is not None:  # final motion vectors from the parsed mvds
            if self._mvf is None:
                from .engine import MotionField
                self._mvf = MotionField(self.width_mbs, self.height_mbs)
            mbs = self._mvf.derive_picture(mbs, self._mvd[self._mb_off[i]:self._mb_off[i + 1]])
        return desc, mbs, levels
