"""TEST INFRASTRUCTURE ONLY: ctypes access to oracle/librefcpu.so (refcpu.c).

Each function restates one step of the reference's reconstruction; refcpu.c cites
the reference file:line it follows for every one of them.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_int, c_int32, c_uint8, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librefcpu.so")

_lib = None


def build() -> str:
    """Compiles librefcpu.so with gcc (the oracle is plain C)."""
    subprocess.run(["make", "-C", HERE, "oracle"], check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.refcpu_decode_picture.restype = c_int
        L.refcpu_decode_picture.argtypes = [c_void_p, c_void_p, c_int, c_int, c_void_p, POINTER(c_void_p), c_int]
        L.refcpu_residual4x4.argtypes = [c_void_p, c_int, c_int, c_int, c_void_p]
        L.refcpu_residual8x8.argtypes = [c_void_p, c_int, c_int, c_void_p]
        L.refcpu_luma_dc.argtypes = [c_void_p, c_int, c_void_p]
        L.refcpu_chroma_dc.argtypes = [c_void_p, c_int, c_void_p]
        L.refcpu_chroma_qp.restype = c_int
        L.refcpu_chroma_qp.argtypes = [c_int, c_int]
        L.refcpu_luma_sample.restype = c_int
        L.refcpu_luma_sample.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int]
        L.refcpu_chroma_sample.restype = c_int
        L.refcpu_chroma_sample.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int]
        L.refcpu_filter_line.argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int]
        L.refcpu_yuv420p_to_bgr24.argtypes = [c_int, c_int, c_void_p, c_void_p, c_int, c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class Dpb:
    """Host DPB of one stream: ``slots`` pictures, each Y then Cb then Cr, followed by a
    zeroed tail of one chroma row that field macroblocks of MBAFF frames read past the
    Cr plane (see refcpu.h)."""

    def __init__(self, width_mbs: int, height_mbs: int, slots: int):
        self.W, self.H = width_mbs, height_mbs
        self.slot_bytes = width_mbs * height_mbs * 384
        self.planes = [np.zeros(self.slot_bytes + width_mbs * 8, np.uint8) for _ in range(slots)]
        self._ptrs = (c_void_p * slots)(*[p.ctypes.data for p in self.planes])

    def decode(self, desc, mbs: np.ndarray, levels: np.ndarray, stage: int = 2) -> int:
        """refcpu_decode_picture on this DPB; returns its status code."""
        lv = np.ascontiguousarray(levels, dtype=np.int16)
        return lib().refcpu_decode_picture(ctypes.addressof(desc), _ptr(mbs), self.W, self.H,
                                           _ptr(lv) if lv.size else None, self._ptrs, stage)

    def split(self, slot: int):
        """(Y, Cb, Cr) views of one slot as 2-D arrays."""
        p = self.planes[slot][:self.slot_bytes]
        ny, nc = self.W * self.H * 256, self.W * self.H * 64
        return (p[:ny].reshape(self.H * 16, self.W * 16), p[ny:ny + nc].reshape(self.H * 8, self.W * 8),
                p[ny + nc:].reshape(self.H * 8, self.W * 8))


def residual4x4(scan, qp: int, dc_pass: int = 0, field: int = 0) -> np.ndarray:
    s = np.ascontiguousarray(scan, np.int32)
    r = np.zeros(16, np.int32)
    lib().refcpu_residual4x4(_ptr(s), qp, dc_pass, field, _ptr(r))
    return r.reshape(4, 4)


def residual8x8(scan, qp: int, field: int = 0) -> np.ndarray:
    s = np.ascontiguousarray(scan, np.int32)
    r = np.zeros(64, np.int32)
    lib().refcpu_residual8x8(_ptr(s), qp, field, _ptr(r))
    return r.reshape(8, 8)


def luma_dc(scan, qp: int) -> np.ndarray:
    s = np.ascontiguousarray(scan, np.int32)
    r = np.zeros(16, np.int32)
    lib().refcpu_luma_dc(_ptr(s), qp, _ptr(r))
    return r.reshape(4, 4)


def chroma_dc(c, qpc: int) -> np.ndarray:
    s = np.ascontiguousarray(c, np.int32)
    r = np.zeros(4, np.int32)
    lib().refcpu_chroma_dc(_ptr(s), qpc, _ptr(r))
    return r


def chroma_qp(qpy: int, offset: int) -> int:
    return lib().refcpu_chroma_qp(qpy, offset)


def luma_sample(ref: np.ndarray, x: int, y: int, xf: int, yf: int) -> int:
    r = np.ascontiguousarray(ref, np.uint8)
    return lib().refcpu_luma_sample(_ptr(r), r.shape[1], r.shape[0], x, y, xf, yf)


def chroma_sample(ref: np.ndarray, x: int, y: int, xf: int, yf: int) -> int:
    r = np.ascontiguousarray(ref, np.uint8)
    return lib().refcpu_chroma_sample(_ptr(r), r.shape[1], r.shape[0], x, y, xf, yf)


def filter_line(p, q, bS: int, chroma: int, qp_avg: int, off_a: int = 0, off_b: int = 0):
    pp = np.ascontiguousarray(p, np.uint8).copy()
    qq = np.ascontiguousarray(q, np.uint8).copy()
    lib().refcpu_filter_line(_ptr(pp), _ptr(qq), bS, chroma, qp_avg, off_a, off_b)
    return pp, qq


def yuv420p_to_bgr24(y: np.ndarray, cb: np.ndarray, cr: np.ndarray, flip: int = 0) -> np.ndarray:
    """BGR24 image (rows padded to 4 bytes, as the reference's BMP output) of one picture."""
    h, w = y.shape
    yuv = np.concatenate([np.ascontiguousarray(p, np.uint8).reshape(-1) for p in (y, cb, cr)])
    stride = (w * 3 + 3) // 4 * 4
    out = np.zeros(stride * h, np.uint8)
    lib().refcpu_yuv420p_to_bgr24(w, h, _ptr(yuv), _ptr(out), stride, flip)
    return out.reshape(h, stride)
