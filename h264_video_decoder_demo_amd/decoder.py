"""Host-side mirror of the reference decoder's public interface.

``H264VideoDecoder`` keeps the shape of CH264VideoDecoder
(h264_video_decoder_demo/H264VideoDecoder.h:22-45): the user registers an output
callback with ``set_output_frame_callback_functuin(callback, user_data)`` and calls
``open(source)``; every decoded picture is handed to
``callback(picture, user_data, error_code)`` in output (POC) order, a non-zero
return stops decoding, and a final ``callback(None, user_data,
H264_DECODE_ERROR_CODE_FILE_END)`` ends the stream (H264VideoDecoder.cpp:340-436).
The output reorder follows CH264PicturesGOP::getOneOutPicture
(H264PicturesGOP.cpp:89-165).

Reconstruction runs on the GPU through libh264g; ``source`` is an H.264 Annex B
byte stream (a ``.h264`` / ``.264`` path, bytes, or a ``bitstream.Bitstream``:
CAVLC frame pictures, parsed on the host by the library's front end), or parsed
per-picture syntax (the generator, or a syntax file written by ``syntax_io``).
"""
from __future__ import annotations

import queue
import threading
import warnings
from dataclasses import dataclass

import numpy as np

from ._lib import PicDesc
from .engine import Engine

# H264_DECODE_ERROR_CODE (H264CommonFunc.h:477-484)
H264_DECODE_ERROR_CODE_UNKNOWN = 0
H264_DECODE_ERROR_CODE_FILE_END = 1
H264_DECODE_ERROR_CODE_OPEN_FILE_FAILED = 2
H264_DECODE_ERROR_CODE_NEED_MORE_DATA = 3
H264_DECODE_ERROR_CODE_NO_MORE_FRAMES = 4

_SLICE_NAMES = {0: "P", 1: "B", 2: "I"}


@dataclass
class H264Picture:
    """The fields of CH264Picture::m_picture_frame that output callbacks read
    (main.cpp:25-31): planar samples with stride = width, sizes and POC."""
    m_pic_buff_luma: np.ndarray
    m_pic_buff_cb: np.ndarray
    m_pic_buff_cr: np.ndarray
    PicWidthInSamplesL: int
    PicHeightInSamplesL: int
    PicOrderCnt: int
    slice_type: int
    m_PicNumCnt: int  # decoding order

    @property
    def slice_type_str(self) -> str:
        return _SLICE_NAMES.get(self.slice_type % 5, "?")

    def yuv420p(self) -> bytes:
        """Y, Cb, Cr planes back to back (the layout of CH264PictureBase's buffer)."""
        return self.m_pic_buff_luma.tobytes() + self.m_pic_buff_cb.tobytes() + self.m_pic_buff_cr.tobytes()


class _OutputReorder:
    """CH264PicturesGOP::getOneOutPicture: holds up to max_num_reorder_frames decoded
    pictures and releases the smallest POC."""

    def __init__(self, max_num_reorder_frames: int):
        self.depth = max_num_reorder_frames
        self.buf = []

    def push(self, pic):
        if self.depth == 0:
            return pic
        if len(self.buf) < self.depth:
            self.buf.append(pic)
            return None
        i = min(range(len(self.buf)), key=lambda k: self.buf[k].PicOrderCnt)
        if pic.PicOrderCnt < self.buf[i].PicOrderCnt:
            return pic
        out, self.buf[i] = self.buf[i], pic
        return out

    def flush_one(self):
        if not self.buf:
            return None
        i = min(range(len(self.buf)), key=lambda k: self.buf[k].PicOrderCnt)
        return self.buf.pop(i)


def _with_last(items):
    """(item, is_last) for each item of an iterator."""
    it = iter(items)
    cur = next(it, None)
    while cur is not None:
        nxt = next(it, None)
        yield cur, nxt is None
        cur = nxt


class _Prefetch:
    """The items of an iterator produced on a thread of their own, up to `depth` ahead of
    the consumer (the byte-stream parser runs in C with the GIL released, so parsing the
    next pictures overlaps the reconstruction of the current one).  An exception of the
    producer is raised at the consumer; close() stops the producer."""

    _END = object()

    def __init__(self, items, depth: int = 4):
        self._q = queue.Queue(maxsize=depth)
        self._stop = threading.Event()
        self._t = threading.Thread(target=self._run, args=(iter(items),), daemon=True)
        self._t.start()

    def _run(self, it):
        try:
            for x in it:
                while not self._stop.is_set():
                    try:
                        self._q.put((True, x), timeout=0.05)
                        break
                    except queue.Full:
                        continue
                if self._stop.is_set():
                    return
            self._q.put((True, self._END))
        except BaseException as e:  # handed to the consumer
            self._q.put((False, e))

    def __iter__(self):
        while True:
            ok, x = self._q.get()
            if not ok:
                raise x
            if x is self._END:
                return
            yield x

    def close(self):
        self._stop.set()
        while self._t.is_alive():
            try:
                self._q.get(timeout=0.05)
            except queue.Empty:
                pass
        self._t.join()


class H264VideoDecoder:
    def __init__(self, engine: Engine | None = None, device: int = 0):
        self._engine = engine
        self._device = device
        self.m_output_frame_callback = None
        self.m_userData = None

    def init(self) -> int:
        return 0

    def unInit(self) -> int:
        return 0

    def set_output_frame_callback_functuin(self, output_frame_callback, userData) -> int:
        self.m_output_frame_callback = output_frame_callback
        self.m_userData = userData
        return 0

    def _emit(self, pic) -> int:
        if self.m_output_frame_callback is None:
            return 0
        return -1 if self.m_output_frame_callback(pic, self.m_userData, 0) != 0 else 0

    def open(self, source) -> int:
        """Decodes every picture of ``source``: an Annex B byte stream (path ending in
        .h264 / .264 / .26l, bytes, or a bitstream.Bitstream), or an object with
        width_mbs, height_mbs, dpb_slots, num_pictures, num_reorder_frames and
        picture(i) -> (desc, mbs, levels), e.g. Generator or syntax_io.SyntaxFile, or a
        path to a syntax file."""
        from .bitstream import Bitstream
        if isinstance(source, (bytes, bytearray)) or (
                isinstance(source, str) and source.lower().endswith((".h264", ".264", ".26l"))):
            try:
                source = Bitstream(source)
            except OSError:
                if self.m_output_frame_callback is not None:
                    return self.m_output_frame_callback(None, self.m_userData, H264_DECODE_ERROR_CODE_OPEN_FILE_FAILED)
                return -1
        elif isinstance(source, str):
            from .syntax_io import SyntaxFile
            try:
                source = SyntaxFile(source)
            except OSError:
                if self.m_output_frame_callback is not None:
                    return self.m_output_frame_callback(None, self.m_userData, H264_DECODE_ERROR_CODE_OPEN_FILE_FAILED)
                return -1
        eng = self._engine or Engine(device=self._device)
        own = self._engine is None
        W, H = source.width_mbs, source.height_mbs
        sid = eng.open_stream(W, H, source.dpb_slots)
        reorder = _OutputReorder(getattr(source, "num_reorder_frames", 0))
        ret = 0
        stopped = False
        byte_stream = isinstance(source, Bitstream)
        prefetch = None
        if byte_stream:
            prefetch = _Prefetch(source.pictures())
            items = _with_last(prefetch)
        else:
            items = ((source.picture(i), False) for i in range(source.num_pictures))
        inflight = None  # (staged picture, index, desc, idr): submitted, not yet read back

        def finish():
            """Waits for the picture in flight, reads it back (host-visible before any
            callback) and hands it to the reorder buffer; True when the callback stops."""
            nonlocal inflight
            pic, i, desc, idr = inflight
            inflight = None
            eng.sync()
            y, cb, cr = eng.readback(sid, desc.dst_slot)
            eng.release(pic)
            if i > 0 and idr:
                while True:  # an IDR flushes the reorder buffer (H264VideoDecoder.cpp:383-410)
                    out = reorder.flush_one()
                    if out is None:
                        break
                    if self._emit(out):
                        return True
            out = reorder.push(H264Picture(y, cb, cr, W * 16, H * 16, desc.poc, desc.slice_type, i))
            return out is not None and self._emit(out)

        try:
            for i, (item, last) in enumerate(items):
                if byte_stream:
                    p = item
                    if p.lost_mbs:
                        # the reference abandons a slice whose data fails and decodes on
                        # (H264VideoDecoder.cpp:135-136); its MBs from there stay zero
                        warnings.warn(f"picture {i} (POC {p.desc.poc}): {p.lost_mbs} MBs from MB {p.first_lost_mb} "
                                      f"not decoded ({p.damage})", RuntimeWarning, stacklevel=2)
                    desc, mbs, levels, idr = p.desc, p.mbs, p.levels, p.idr
                else:
                    (desc, mbs, levels), idr = item, source.is_idr(i)
                if last and byte_stream:
                    # the reference hands the stream's last picture to the callback without
                    # finishing it (H264VideoDecoder.cpp:353-361 skips
                    # end_decode_the_picture_and_get_a_new_empty_picture, which runs the
                    # deblocking filter, H264PictureBase.cpp:707): it is output unfiltered
                    desc = PicDesc.from_buffer_copy(desc)
                    desc.disable_deblocking_filter_idc = 1
                # staging this picture on the host overlaps the reconstruction of the one in
                # flight; it is submitted once that one is read back, since a picture may
                # write the slot the previous one was output from
                pic = eng.begin(sid, desc, mbs, levels)
                if inflight is not None and finish():
                    eng.release(pic)
                    stopped = True
                    break
                eng.submit(pic)
                inflight = (pic, i, desc, idr)
            if not stopped and inflight is not None:
                stopped = finish()
            if not stopped:
                while True:
                    out = reorder.flush_one()
                    if out is None:
                        break
                    if self._emit(out):
                        stopped = True
                        break
        finally:
            if inflight is not None:
                eng.sync()
                eng.release(inflight[0])
            if prefetch is not None:
                prefetch.close()
            eng.close_stream(sid)
            if own:
                eng.close()
        if self.m_output_frame_callback is not None:
            ret = self.m_output_frame_callback(None, self.m_userData, H264_DECODE_ERROR_CODE_FILE_END)
        return ret
