"""MI355X (gfx950) H.264 picture-reconstruction engine.

A drop-in for the data-parallel half of jfu222/h264_video_decoder_demo: dequant and
inverse transforms, inter prediction with weighted prediction, intra prediction and
deblocking run as hand-written HIP kernels behind the C ABI of include/h264g.h;
parsing stays on the host.  ``decoder.H264VideoDecoder`` mirrors the reference's
decode/output-callback interface (H264VideoDecoder.h).
"""
from .engine import Engine, Generator, MotionField, StagedPicture, decode_stream_plain  # noqa: F401
from .decoder import H264VideoDecoder, H264Picture, H264_DECODE_ERROR_CODE_FILE_END  # noqa: F401
