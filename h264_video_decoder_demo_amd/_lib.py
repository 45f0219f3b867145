"""ctypes binding of libh264g (include/h264g.h).

The library is built in-tree by ``__graft_entry__.build()`` (hipcc for gfx950).
Loading fails loudly when it is missing: there is no CPU fallback for the engine.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int8, c_int16, c_int32, c_uint8, c_uint16, c_uint32, c_uint64, c_void_p

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libh264g.so")

# error codes
H264G_OK = 0
H264G_E_ARG = -1
H264G_E_NODEV = -2
H264G_E_NOMEM = -3
H264G_E_BUDGET = -4
H264G_E_STATE = -5
H264G_E_DEVICE = -6
H264G_E_SYNTAX = -7
H264G_E_TIMEOUT = -8

SLICE_P, SLICE_B, SLICE_I = 0, 1, 2
MB_I4x4, MB_I8x8, MB_I16x16, MB_IPCM, MB_P, MB_P_SKIP, MB_B = range(7)
MBF_T8x8, MBF_FIELD, MBF_SKIP, MBF_LOST, MBF_DIRECT_ALL = 0x01, 0x02, 0x04, 0x08, 0xF0
SLICE_NONE = 0xFFFFFFFF


class H264GError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"h264g error {code}: {msg}")
        self.code = code


class MbRecord(ctypes.Structure):
    _fields_ = [
        ("mb_type", c_uint8), ("flags", c_uint8), ("qp_y", c_uint8), ("cbp", c_uint8),
        ("intra16x16_pred_mode", c_uint8), ("intra_chroma_pred_mode", c_uint8), ("part_shape", c_uint8),
        ("reserved0", c_uint8), ("sub_part_shape", c_uint8 * 4), ("intra_pred_modes", c_uint8 * 16),
        ("ref_idx", (c_int8 * 4) * 2), ("mv", ((c_int16 * 2) * 16) * 2), ("level_offset", c_uint32),
        ("slice_num", c_uint32),
    ]


MB_DTYPE = np.dtype([
    ("mb_type", "u1"), ("flags", "u1"), ("qp_y", "u1"), ("cbp", "u1"),
    ("intra16x16_pred_mode", "u1"), ("intra_chroma_pred_mode", "u1"), ("part_shape", "u1"),
    ("reserved0", "u1"), ("sub_part_shape", "u1", (4,)), ("intra_pred_modes", "u1", (16,)),
    ("ref_idx", "i1", (2, 4)), ("mv", "<i2", (2, 16, 2)), ("level_offset", "<u4"), ("slice_num", "<u4"),
])
assert MB_DTYPE.itemsize == ctypes.sizeof(MbRecord) == 172


class PicDesc(ctypes.Structure):
    _fields_ = [
        ("slice_type", c_int32), ("structure", c_int32), ("mbaff", c_int32),
        ("constrained_intra_pred_flag", c_int32), ("disable_deblocking_filter_idc", c_int32),
        ("filter_offset_a", c_int32), ("filter_offset_b", c_int32),
        ("chroma_qp_index_offset", c_int32), ("second_chroma_qp_index_offset", c_int32),
        ("weighted_pred_flag", c_int32), ("weighted_bipred_idc", c_int32),
        ("luma_log2_weight_denom", c_int32), ("chroma_log2_weight_denom", c_int32),
        ("num_ref_idx_active", c_int32 * 2), ("poc", c_int32), ("dst_slot", c_int32),
        ("ref_slot", (c_int32 * 32) * 2), ("ref_poc", (c_int32 * 32) * 2), ("ref_long_term", (c_int32 * 32) * 2),
        ("luma_weight_flag_last", c_int32 * 2), ("luma_weight", (c_int16 * 32) * 2), ("luma_offset", (c_int16 * 32) * 2),
        ("chroma_weight", ((c_int16 * 2) * 32) * 2), ("chroma_offset", ((c_int16 * 2) * 32) * 2),
        ("scaling_matrix", c_int32), ("weight_scale4x4", (c_uint8 * 16) * 6), ("weight_scale8x8", (c_uint8 * 64) * 2),
    ]


class StreamDesc(ctypes.Structure):
    _fields_ = [
        ("pic_width_in_mbs", c_int32), ("frame_height_in_mbs", c_int32), ("chroma_array_type", c_int32),
        ("bit_depth", c_int32), ("frame_mbs_only", c_int32), ("mbaff", c_int32), ("dpb_slots", c_int32),
    ]


class Config(ctypes.Structure):
    _fields_ = [
        ("device", c_int32), ("dpb_budget_bytes", c_uint64), ("max_batch_pictures", c_int32),
        ("profile", c_int32), ("waves", c_int32), ("no_graphs", c_int32), ("deblock_schedule", c_int32),
        ("deblock_waves", c_int32), ("intra_schedule", c_int32), ("intra_waves", c_int32),
        ("lane_split_mbs", c_int32),
    ]


DEBLOCK_BANDS, DEBLOCK_MB_WAVEFRONT, DEBLOCK_MB_ROWS = 0, 1, 2
INTRA_BANDS, INTRA_MB_WAVEFRONT = 0, 1


class GenParams(ctypes.Structure):
    _fields_ = [
        ("config", c_int32), ("seed", c_uint64), ("width_mbs", c_int32), ("height_mbs", c_int32),
        ("n_frames", c_int32), ("gop", c_int32), ("b_frames", c_int32), ("max_refs", c_int32), ("t8x8", c_int32),
        ("mv_mode", c_int32), ("mv_range", c_int32), ("mv_clamp_lo", c_int32), ("mv_clamp_hi", c_int32),
        ("qp_lo", c_int32), ("qp_hi", c_int32), ("zero_permille", c_int32), ("laplace_b_x100", c_int32),
        ("intra_in_inter_permille", c_int32), ("skip_permille", c_int32),
        ("weighted_pred_flag", c_int32), ("weighted_bipred_idc", c_int32), ("filter_offsets", c_int32),
        ("disable_deblocking_filter_idc", c_int32), ("chroma_qp_index_offset", c_int32),
        ("second_chroma_qp_index_offset", c_int32), ("ipcm_permille", c_int32), ("p_pattern", c_int32),
        ("spare_slots", c_int32), ("mbaff", c_int32), ("constrained_intra_pred", c_int32),
        ("direct_permille", c_int32), ("slices", c_int32),
    ]


class ParsedPicture(ctypes.Structure):
    _fields_ = [
        ("desc", PicDesc), ("mbs", c_void_p), ("n_mbs", c_uint32), ("levels", c_void_p), ("n_levels", c_uint64),
        ("width_mbs", c_int32), ("height_mbs", c_int32), ("dpb_slots", c_int32), ("max_num_reorder_frames", c_int32),
        ("idr", c_int32), ("nal_ref_idc", c_int32), ("frame_num", c_int32),
        ("profile_idc", c_int32), ("constraint_set3", c_int32), ("level_idc", c_int32),
        ("n_slices", c_int32), ("lost_mbs", c_int32), ("first_lost_mb", c_int32),
    ]


# Every entry point include/h264g.h declares, with its ctypes signature.
_SIGNATURES = {
    "h264g_create": (c_int32, [POINTER(Config), POINTER(c_void_p)]),
    "h264g_destroy": (c_int32, [c_void_p]),
    "h264g_stream_open": (c_int32, [c_void_p, POINTER(StreamDesc), POINTER(c_uint32)]),
    "h264g_stream_close": (c_int32, [c_void_p, c_uint32]),
    "h264g_set_bands": (c_int32, [c_void_p, c_int32, c_int32]),
    "h264g_picture_begin": (c_int32, [c_void_p, c_uint32, POINTER(PicDesc), c_void_p, c_uint32, c_void_p, c_uint64,
                                      POINTER(c_void_p)]),
    "h264g_picture_submit": (c_int32, [c_void_p, c_void_p]),
    "h264g_picture_decode_static": (c_int32, [c_void_p, c_uint32, POINTER(PicDesc), c_void_p, c_uint32, c_void_p,
                                              c_uint64]),
    "h264g_batch_submit": (c_int32, [c_void_p, POINTER(c_void_p), c_uint32]),
    "h264g_batch_create": (c_int32, [c_void_p, POINTER(c_void_p), c_uint32, POINTER(c_void_p)]),
    "h264g_batch_run": (c_int32, [c_void_p, c_void_p]),
    "h264g_batch_destroy": (c_int32, [c_void_p, c_void_p]),
    "h264g_stage_times": (c_int32, [c_void_p, POINTER(c_float)]),
    "h264g_set_profile": (c_int32, [c_void_p, c_int32]),
    "h264g_fence_record": (c_int32, [c_void_p, POINTER(c_void_p)]),
    "h264g_fence_query": (c_int32, [c_void_p, c_void_p]),
    "h264g_fence_wait": (c_int32, [c_void_p, c_void_p]),
    "h264g_fence_destroy": (c_int32, [c_void_p, c_void_p]),
    "h264g_sync": (c_int32, [c_void_p]),
    "h264g_slot_readback": (c_int32, [c_void_p, c_uint32, c_int32, c_void_p, c_void_p, c_void_p]),
    "h264g_slot_upload": (c_int32, [c_void_p, c_uint32, c_int32, c_void_p, c_void_p, c_void_p]),
    "h264g_slot_to_bgr24": (c_int32, [c_void_p, c_uint32, c_int32, c_void_p, c_uint32, c_int32]),
    "h264g_picture_release": (c_int32, [c_void_p, c_void_p]),
    "h264g_picture_alg_bytes": (c_int32, [c_void_p, POINTER(c_uint64)]),
    "h264g_calibration_copy": (c_int32, [c_void_p, c_uint64, c_int32]),
    "h264g_last_error": (c_char_p, []),
    "h264g_device_describe": (c_int32, [c_int32, c_char_p, c_uint32]),
    "h264g_gen_default_params": (c_int32, [c_int32, POINTER(GenParams)]),
    "h264g_gen_create": (c_int32, [POINTER(GenParams), POINTER(c_void_p)]),
    "h264g_gen_destroy": (c_int32, [c_void_p]),
    "h264g_gen_num_pictures": (c_int32, [c_void_p]),
    "h264g_gen_picture": (c_int32, [c_void_p, c_int32, POINTER(PicDesc), POINTER(c_void_p), POINTER(c_uint32),
                                    POINTER(c_void_p), POINTER(c_uint64)]),
    "h264g_gen_dpb_slots": (c_int32, [c_void_p]),
    "h264g_gen_digest": (c_uint64, [c_void_p]),
    "h264g_mvfield_create": (c_int32, [c_int32, c_int32, POINTER(c_void_p)]),
    "h264g_mvfield_destroy": (None, [c_void_p]),
    "h264g_mvfield_reset": (c_int32, [c_void_p]),
    "h264g_mvfield_derive": (c_int32, [c_void_p, c_uint32, c_void_p, c_void_p]),
    "h264g_mvfield_intra_modes": (c_int32, [c_void_p, c_uint32, c_void_p, c_void_p, c_int32]),
    "h264g_mvfield_colocated": (c_int32, [c_void_p, c_void_p, c_int32]),
    "h264g_parser_create": (c_int32, [c_void_p, c_uint64, POINTER(c_void_p)]),
    "h264g_parser_destroy": (c_int32, [c_void_p]),
    "h264g_parser_next": (c_int32, [c_void_p, POINTER(ParsedPicture)]),
    "h264g_parser_damage": (ctypes.c_char_p, [c_void_p]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Loads libh264g.so (raises when the HIP extension has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc < 0:
        raise H264GError(rc, lib().h264g_last_error().decode(errors="replace"))
    return rc
