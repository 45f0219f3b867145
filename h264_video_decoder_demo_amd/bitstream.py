"""H.264 Annex B byte streams as a picture source for the decoder and the engine.

``Bitstream`` wraps the library's front end (h264g_parser_*: NAL units, parameter
sets, slice headers, CAVLC slice data, POC, reference marking and lists, DPB slots)
and yields each picture in decoding order as (desc, mbs, levels) ready for
``Engine.begin``, with the stream properties the output reorder needs.  It stands in
for the reference's own parsing on the way to reconstruction
(H264VideoDecoder.cpp:56-380, H264SliceData.cpp, H264MacroBlock.cpp,
H264ResidualBlockCavlc.cpp, H264RefPicList.cpp).
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_void_p
from dataclasses import dataclass

import numpy as np

from ._lib import MB_DTYPE, ParsedPicture, PicDesc, check, lib

# Table A-1: (level_idc, MaxDpbMbs), for the inferred max_num_reorder_frames
_MAX_DPB_MBS = {9: 396, 10: 396, 11: 900, 12: 2376, 13: 2376, 20: 2376, 21: 4752, 22: 8100, 30: 8100, 31: 18000,
                32: 20480, 40: 32768, 41: 32768, 42: 34816, 50: 110400, 51: 184320, 52: 184320}


@dataclass
class Picture:
    desc: PicDesc
    mbs: np.ndarray
    levels: np.ndarray
    idr: bool
    nal_ref_idc: int
    frame_num: int
    n_slices: int = 1
    lost_mbs: int = 0         # MBs no slice decoded (a damaged slice stopped early): zero samples
    first_lost_mb: int = -1
    damage: str = ""          # the parse error that stopped the damaged slice


class Bitstream:
    """Pictures of an Annex B byte stream (a path or bytes), parsed on the host."""

    def __init__(self, source):
        if isinstance(source, (bytes, bytearray, memoryview)):
            data = bytes(source)
        else:
            with open(source, "rb") as f:
                data = f.read()
        self._data = data
        self._h = None
        self._first = None
        self._restart()
        self._first = self._next()
        if self._first is None:
            raise ValueError("no picture in the stream")
        pp = self._pp
        self.width_mbs, self.height_mbs, self.dpb_slots = pp.width_mbs, pp.height_mbs, pp.dpb_slots
        self.num_reorder_frames = self._reorder_depth(pp)

    @staticmethod
    def _reorder_depth(pp) -> int:
        """max_num_reorder_frames as the reference takes it (H264SPS.cpp:365-396): the VUI
        value, else 0 for the intra profiles with constraint_set3_flag, else MaxDpbFrames of
        the level (Table A-1, at most 16)."""
        if pp.max_num_reorder_frames >= 0:
            return pp.max_num_reorder_frames
        if pp.profile_idc in (44, 86, 100, 110, 122, 244) and pp.constraint_set3:
            return 0
        mbs = _MAX_DPB_MBS.get(pp.level_idc)
        return min(mbs // (pp.width_mbs * pp.height_mbs), 16) if mbs else 0

    def _restart(self):
        if self._h is not None:
            lib().h264g_parser_destroy(self._h)
        self._h = c_void_p()
        buf = ctypes.create_string_buffer(self._data, len(self._data))
        check(lib().h264g_parser_create(buf, len(self._data), byref(self._h)))
        self._pp = ParsedPicture()

    def _next(self):
        rc = check(lib().h264g_parser_next(self._h, byref(self._pp)))
        if rc == 0:
            return None
        pp = self._pp
        mbs = np.ctypeslib.as_array(ctypes.cast(pp.mbs, ctypes.POINTER(ctypes.c_uint8)),
                                    shape=(pp.n_mbs * MB_DTYPE.itemsize,)).view(MB_DTYPE).copy()
        lv = (np.ctypeslib.as_array(ctypes.cast(pp.levels, ctypes.POINTER(ctypes.c_int16)), shape=(pp.n_levels,)).copy()
              if pp.n_levels else np.zeros(0, np.int16))
        desc = PicDesc.from_buffer_copy(pp.desc)
        damage = lib().h264g_parser_damage(self._h).decode() if pp.lost_mbs else ""
        return Picture(desc, mbs, lv, bool(pp.idr), pp.nal_ref_idc, pp.frame_num, pp.n_slices, pp.lost_mbs,
                       pp.first_lost_mb, damage)

    def pictures(self):
        """Every picture in decoding order (parses the stream from its start)."""
        if self._first is not None:
            first, self._first = self._first, None
        else:
            self._restart()
            first = self._next()
        p = first
        while p is not None:
            yield p
            p = self._next()

    def close(self):
        if self._h is not None:
            lib().h264g_parser_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
