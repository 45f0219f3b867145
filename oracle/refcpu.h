/*
 * refcpu.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference decoder's picture reconstruction
 * (jfu222/h264_video_decoder_demo) used as the oracle for the HIP engine: only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg call it, as the
 * checker, never as the thing measured or shipped.
 */
#ifndef REFCPU_H
#define REFCPU_H

#include <stdint.h>
#include "h264g.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Reconstructs one picture (a frame or an MBAFF frame, one slice) into
 * planes[desc->dst_slot].  planes[s] points at slot s: Y (16W x 16H) followed by Cb and
 * Cr (8W x 8H each), stride = width, then 8W zero bytes: field macroblocks of MBAFF
 * frames read up to that far past the Cr plane, as the reference does past its frame
 * buffer (quirk Q17 in refcpu.c).  stage: 1 = stop before deblocking, 2 = full
 * reconstruction.  Returns 0, or a negative H264G_E_* code for syntax the reference
 * rejects.
 */
int refcpu_decode_picture(const h264g_pic_desc *desc, const h264g_mb *mbs, int width_mbs, int height_mbs,
                          const int16_t *levels, uint8_t *const *planes, int stage);

/* ---- single-function entry points (known-answer and property tests) ---- */
/* 8.5.12: dequantise + inverse 4x4 transform of one block given in scan order. */
void refcpu_residual4x4(const int32_t scan[16], int qp, int dc_passthrough, int field, int32_t r[16]);
/* 8.5.13: dequantise + inverse 8x8 transform. */
void refcpu_residual8x8(const int32_t scan[64], int qp, int field, int32_t r[64]);
/* 8.5.10: Intra16x16 luma DC (scan order in, raster dcY[4][4] out). */
void refcpu_luma_dc(const int32_t scan[16], int qp, int32_t dcY[16]);
/* 8.5.11: 4:2:0 chroma DC (c[0..3] raster 2x2 in, dcC out). */
void refcpu_chroma_dc(const int32_t c[4], int qpc, int32_t dcC[4]);
/* 8.5.8 / Table 8-15 */
int refcpu_chroma_qp(int qpy, int offset);
/* 8.4.2.2.1: one luma sample of ref (w x h, clamped coordinates). */
int refcpu_luma_sample(const uint8_t *ref, int w, int h, int xInt, int yInt, int xFrac, int yFrac);
/* 8.4.2.2.2: one chroma sample. */
int refcpu_chroma_sample(const uint8_t *ref, int w, int h, int xInt, int yInt, int xFrac, int yFrac);
/* 8.7.2.3 / 8.7.2.4 on one line of 8 samples p3..p0 q0..q3 (in/out). */
void refcpu_filter_line(uint8_t p[4], uint8_t q[4], int bS, int chroma, int qp_avg, int offset_a, int offset_b);

/* YUV420P -> BGR24 of a decoded picture (CH264PictureBase::convertYuv420pToBgr24,
 * H264PictureBase.cpp:440-466): bgr rows of `stride` bytes; flip = 1 writes rows
 * bottom-up (convertYuv420pToBgr24FlipLines). */
void refcpu_yuv420p_to_bgr24(int width, int height, const uint8_t *yuv, uint8_t *bgr, int stride, int flip);

#ifdef __cplusplus
}
#endif

#endif
