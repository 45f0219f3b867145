/*
 * h264g.h -- C ABI of the MI355X (gfx950) H.264 picture-reconstruction engine.
 *
 * The engine owns the data-parallel half of an H.264 decoder: dequantisation and
 * inverse transforms (clause 8.5), inter prediction with fractional-sample
 * interpolation and weighted prediction (8.4.2/8.4.3), intra prediction (8.3)
 * and the deblocking filter (8.7).  Parsing and entropy decoding stay on the
 * host, which hands per-picture syntax arrays to the engine through this ABI.
 *
 * Where the reference decoder (jfu222/h264_video_decoder_demo) performs the same
 * work, each entry point names the reference interface it replaces:
 *
 *   per-MB reconstruction dispatch   h264_video_decoder_demo/H264SliceData.cpp:386-487
 *   P_Skip / B_Skip reconstruction   h264_video_decoder_demo/H264SliceData.cpp:222-227,333-338
 *   per-picture deblocking           h264_video_decoder_demo/H264PictureBase.cpp:707
 *   DPB picture storage              h264_video_decoder_demo/H264PictureBase.cpp:160-182
 *   output callback (samples read)   h264_video_decoder_demo/H264VideoDecoder.h:22, main.cpp:25-31
 *
 * Conventions (same as the reference, h264_video_decoder_demo/CommonFunction.h:40-48):
 * every function returns int32: 0 = success, negative = error (H264G_E_*).
 * Handles are opaque.  Input arrays are copied before the call returns, so the
 * caller keeps ownership.  Calls on one stream must be serialised by the caller.
 * All structures are plain C with fixed-width fields; no torch types cross the ABI.
 */
#ifndef H264G_H
#define H264G_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ errors */
#define H264G_OK             0
#define H264G_E_ARG         -1   /* invalid argument / malformed syntax record      */
#define H264G_E_NODEV       -2   /* no usable gfx950 device                          */
#define H264G_E_NOMEM       -3   /* host or device allocation failed                 */
#define H264G_E_BUDGET      -4   /* stream admission refused: DPB HBM budget        */
#define H264G_E_STATE       -5   /* call not valid in the current state              */
#define H264G_E_DEVICE      -6   /* a HIP runtime call failed                        */
#define H264G_E_SYNTAX      -7   /* syntax that the reference rejects (e.g. WP sum)  */
#define H264G_E_TIMEOUT     -8   /* a device wavefront wait exceeded its spin limit  */

/* ----------------------------------------------------------- enumerations */
/* slice_type % 5, as in h264_video_decoder_demo/H264CommonFunc.h (P=0, B=1, I=2). */
#define H264G_SLICE_P   0
#define H264G_SLICE_B   1
#define H264G_SLICE_I   2

/* picture structure (h264_video_decoder_demo/H264CommonFunc.h H264_PICTURE_CODED_TYPE) */
#define H264G_FRAME         0
#define H264G_TOP_FIELD     1
#define H264G_BOTTOM_FIELD  2

/* Macroblock class.  The exact mb_type name follows from the class, the partition
 * shape and which lists each partition uses (ref_idx >= 0), the way
 * CH264MacroBlock::MbPartPredMode resolves it (H264MacroBlock.cpp:509). */
#define H264G_MB_I4x4     0   /* I_NxN, transform_size_8x8_flag = 0 */
#define H264G_MB_I8x8     1   /* I_NxN, transform_size_8x8_flag = 1 */
#define H264G_MB_I16x16   2   /* I_16x16_<mode>_<cbp>               */
#define H264G_MB_IPCM     3   /* I_PCM                              */
#define H264G_MB_P        4   /* P_L0_16x16 / P_L0_L0_16x8 / P_L0_L0_8x16 / P_8x8 */
#define H264G_MB_P_SKIP   5   /* P_Skip (final MV supplied by the host)        */
#define H264G_MB_B        6   /* B_{L0,L1,Bi}_*, B_8x8; with the DIRECT flags also
                                 B_Skip, B_Direct_16x16 and B_Direct_8x8 sub-MBs  */

/* Partition shapes (MbPartWidth x MbPartHeight) and sub-partition shapes. */
#define H264G_PART_16x16  0
#define H264G_PART_16x8   1
#define H264G_PART_8x16   2
#define H264G_PART_8x8    3
#define H264G_SUB_8x8     0
#define H264G_SUB_8x4     1
#define H264G_SUB_4x8     2
#define H264G_SUB_4x4     3

/* h264g_mb.flags */
#define H264G_MBF_T8x8    0x01u   /* transform_size_8x8_flag                 */
#define H264G_MBF_FIELD   0x02u   /* mb_field_decoding_flag (MBAFF pairs)    */
/* Direct prediction (8.4.1.2, direct_8x8_inference_flag = 1) in B pictures: an H264G_MB_B
 * record with part_shape 8x8 whose quadrant q is a direct-predicted 8x8 block carries
 * H264G_MBF_DIRECT(q), sub_part_shape[q] = H264G_SUB_8x8 and the derived ref_idx / mv of
 * the quadrant (direct motion is uniform over each 8x8 quadrant under the inference).
 * All four bits: B_Direct_16x16, or with H264G_MBF_SKIP B_Skip (no residual, cbp 0,
 * QPY = QPY,PREV).  Some bits: B_8x8 with B_Direct_8x8 sub-macroblocks.  The device path
 * reads only ref_idx and mv; the host derivation (h264g_mvfield_*) fills them. */
#define H264G_MBF_SKIP    0x04u
/* Marks (for the caller; reconstruction reads the record as it stands) an MB of a
 * damaged slice: the MB whose slice data failed to parse, reconstructed from the syntax
 * that parsed before the failure with QPY 0 (the reference's reset value), or, when no
 * syntax of it is usable, an I_PCM of zero samples; and every MB that no slice reached, an
 * I_PCM of zero samples with QPY 0 and slice_num H264G_SLICE_NONE, as the reference leaves
 * such MBs as each picture's reset zeroed them (H264PictureBase.cpp:45-67).
 * Deblocking stops at the first MB (address order) whose slice_num is H264G_SLICE_NONE:
 * that MB and every later one filter none of their edges, as the reference's bS
 * derivation fails on the reset MB type (H264InterPrediction.cpp:3182-3186) and the
 * failure ends the picture's deblocking (H264PictureBase.cpp:707-708). */
#define H264G_MBF_LOST    0x08u
#define H264G_SLICE_NONE  0xFFFFFFFFu
#define H264G_MBF_DIRECT(q) (0x10u << (q))
#define H264G_MBF_DIRECT_ALL 0xF0u

/* ----------------------------------------------------------------- records */
/*
 * One fixed-size record per macroblock, in macroblock-address order.  The fields
 * are the ones CH264MacroBlock carries into reconstruction after parsing
 * (H264MacroBlock.h:113-223): types, QPY, CBP, resolved intra prediction modes,
 * reference indices and the final motion vectors (mvp + mvd, or the P_Skip
 * prediction), plus where the MB's coefficient levels start.
 *
 * mv[list][blk][c]: quarter-sample MV of the 4x4 luma block blk = 4*y4 + x4
 * (raster order inside the MB), component c (0 = x, 1 = y).  Every 4x4 block of a
 * partition carries the partition's MV.  Unused lists are ignored.
 *
 * ref_idx[list][q]: reference index of the 8x8 quadrant q (raster order) in
 * RefPicList<list>, or -1 when the quadrant does not predict from that list.
 */
typedef struct h264g_mb {
    uint8_t  mb_type;                 /* H264G_MB_*                                  */
    uint8_t  flags;                   /* H264G_MBF_*                                 */
    uint8_t  qp_y;                    /* QPY, 0..51                                  */
    uint8_t  cbp;                     /* CodedBlockPatternLuma | Chroma << 4         */
    uint8_t  intra16x16_pred_mode;    /* Intra16x16PredMode, 0..3                    */
    uint8_t  intra_chroma_pred_mode;  /* intra_chroma_pred_mode, 0..3                */
    uint8_t  part_shape;              /* H264G_PART_*   (inter)                      */
    uint8_t  reserved0;
    uint8_t  sub_part_shape[4];       /* H264G_SUB_*    (part_shape == 8x8)          */
    uint8_t  intra_pred_modes[16];    /* Intra4x4PredMode[luma4x4BlkIdx] or
                                         Intra8x8PredMode[luma8x8BlkIdx] (first 4)  */
    int8_t   ref_idx[2][4];
    int16_t  mv[2][16][2];
    uint32_t level_offset;            /* first int16 of this MB's level block        */
    uint32_t slice_num;               /* the MB's slice: neighbours with another value
                                         are not available for intra prediction and
                                         motion vector prediction; their common edge
                                         is not filtered with
                                         disable_deblocking_filter_idc 2, and with 0
                                         it is filtered as the reference does, with
                                         the current MB in the neighbour's place for
                                         bS and qP (0 for every MB of a picture of
                                         one slice)                                  */
} h264g_mb;

/*
 * Coefficient levels of one macroblock, starting at levels[level_offset]:
 *
 *   uint16 masks[n_masks]   then   int16 values[sum of popcount(masks)]
 *
 * Each mask covers one transform block; bit k set means the level at scan position
 * k (zig-zag, or field scan for field MBs / field pictures) is non-zero, and the
 * values follow in mask order.  Blocks appear in the order residual() parses them
 * (H264MacroBlock.cpp:1592-1856):
 *   1. Intra16x16DCLevel                       (one mask)       if I16x16
 *   2. for i8x8 in 0..3 with CodedBlockPatternLuma bit i8x8:
 *        transform 8x8: one 64-entry block     (four masks, scan positions 0..63)
 *        otherwise: luma4x4BlkIdx 4*i8x8+0..3 (one mask each; for I16x16 the AC
 *        levels sit at scan positions 1..15, bit 0 is zero)
 *   3. ChromaDCLevel Cb, Cr                    (one mask each, bits 0..3) if cbp>>4 >= 1
 *   4. ChromaACLevel Cb 0..3, Cr 0..3          (one mask each, positions 1..15) if cbp>>4 == 2
 * All masks of the MB come first, then all values.  I_PCM MBs instead carry
 * 384 samples (256 Y, 64 Cb, 64 Cr) as int16 with no masks.
 */

/* ---------------------------------------------------------- descriptors */
typedef struct h264g_config {
    int32_t  device;              /* HIP device ordinal                               */
    uint64_t dpb_budget_bytes;    /* HBM budget for all streams' DPB slots (0 = none) */
    int32_t  max_batch_pictures;  /* pictures per batched launch (0 = default 4096)   */
    int32_t  profile;             /* 1 = record per-stage HIP events                  */
    int32_t  waves;               /* persistent waves per wavefront launch (0 = auto) */
    int32_t  no_graphs;           /* 1 = launch batch stages directly; 0 = replay each
                                     batch's stage sequence as a captured hipGraph
                                     (profiled contexts always launch directly)       */
    int32_t  deblock_schedule;    /* H264G_DEBLOCK_*                                  */
    int32_t  deblock_waves;       /* MB rows per band of H264G_DEBLOCK_BANDS, 1..24
                                     (0 = 8); two rows per wave                      */
    int32_t  intra_schedule;      /* H264G_INTRA_*                                    */
    int32_t  intra_waves;         /* rows per band = waves per workgroup of
                                     H264G_INTRA_BANDS, 1..8 (0 = 8)                 */
    int32_t  lane_split_mbs;      /* streams of at least this many MBs per picture run
                                     on a second lane, so their long intra /
                                     deblocking chains overlap the other pictures'
                                     stages (0 = 16384); negative = one HIP stream for
                                     everything, every picture after the previous one
                                     (no deblocking overlap either)                   */
} h264g_config;

/* Intra schedules: BANDS = pictures cut into bands of `intra_waves` MB rows, one
 * workgroup per band and one wave per row walking the row's intra MBs, LDS row progress
 * inside a band and device-memory progress between bands (default); MB_WAVEFRONT =
 * persistent waves over the intra MBs of all pictures ordered by t = x + 2y, progress
 * words in device memory. */
#define H264G_INTRA_BANDS        0
#define H264G_INTRA_MB_WAVEFRONT 1

/* Deblocking schedules (all give the raster-order result of 8.7):
 * BANDS        pictures cut into bands of `deblock_waves` MB rows, one workgroup per
 *              band and two rows per wave (the second one MB behind the first, on
 *              the wave's other 32 lanes); each MB's bottom rows go to the next row
 *              through LDS inside a band and through device memory between bands
 *              (default)
 * MB_WAVEFRONT persistent waves over single MBs ordered by t = x + 2y, progress words
 *              in device memory
 * MB_ROWS      persistent waves over MB rows of all pictures, progress in device memory */
#define H264G_DEBLOCK_BANDS        0
#define H264G_DEBLOCK_MB_WAVEFRONT 1
#define H264G_DEBLOCK_MB_ROWS      2

typedef struct h264g_stream_desc {
    int32_t  pic_width_in_mbs;    /* PicWidthInMbs                                    */
    int32_t  frame_height_in_mbs; /* FrameHeightInMbs                                 */
    int32_t  chroma_array_type;   /* 1 (4:2:0) -- the only format this engine accepts */
    int32_t  bit_depth;           /* 8                                                */
    int32_t  frame_mbs_only;      /* frame_mbs_only_flag                              */
    int32_t  mbaff;               /* mb_adaptive_frame_field_flag                     */
    int32_t  dpb_slots;           /* DPB pictures to reserve, 1..17                   */
} h264g_stream_desc;

/*
 * Per-picture state the reconstruction reads from the slice header, PPS and the
 * reference lists (one slice per picture).  Weighted-prediction syntax is passed
 * raw, as pred_weight_table() leaves it in CH264SliceHeader
 * (H264SliceHeader.cpp:586-665), so the engine derives the per-partition weights
 * itself with the reference's rules (H264InterPrediction.cpp:2833-3050).
 */
typedef struct h264g_pic_desc {
    int32_t  slice_type;                      /* H264G_SLICE_*                         */
    int32_t  structure;                       /* H264G_FRAME; field pictures (_TOP_FIELD / _BOTTOM_FIELD)
                                                 are refused with H264G_E_ARG: the reference decoder
                                                 has no defined output for them */
    int32_t  mbaff;                           /* MbaffFrameFlag                        */
    int32_t  constrained_intra_pred_flag;
    int32_t  disable_deblocking_filter_idc;
    int32_t  filter_offset_a;                 /* slice_alpha_c0_offset_div2 << 1       */
    int32_t  filter_offset_b;                 /* slice_beta_offset_div2 << 1           */
    int32_t  chroma_qp_index_offset;
    int32_t  second_chroma_qp_index_offset;
    int32_t  weighted_pred_flag;
    int32_t  weighted_bipred_idc;
    int32_t  luma_log2_weight_denom;
    int32_t  chroma_log2_weight_denom;
    int32_t  num_ref_idx_active[2];           /* RefPicListXLength                     */
    int32_t  poc;                             /* PicOrderCnt of the current picture    */
    int32_t  dst_slot;                        /* DPB slot that receives the picture    */
    int32_t  ref_slot[2][32];                 /* RefPicListX[i] -> DPB slot            */
    int32_t  ref_poc[2][32];                  /* PicOrderCnt(RefPicListX[i])           */
    int32_t  ref_long_term[2][32];            /* marked "used for long-term reference" */
    int32_t  luma_weight_flag_last[2];        /* last parsed luma_weight_lX_flag       */
    int16_t  luma_weight[2][32];
    int16_t  luma_offset[2][32];
    int16_t  chroma_weight[2][32][2];
    int16_t  chroma_offset[2][32][2];
    /* Scaling matrices (7.4.2.1.1.1, Table 7-2), the lists the reference's
     * set_scaling_lists_values leaves for the picture (H264SliceHeader.cpp:928-1158),
     * in raster order (row * 4 + col, row * 8 + col: the frame inverse scan of 8.5.6 /
     * 8.5.7).  0 in scaling_matrix: flat lists (16), the arrays are not read.  As in the
     * reference, dequantisation uses only the luma lists: chroma blocks take the
     * LevelScale4x4 of the MB's luma list (quirk Q13, H264PictureBase.cpp:3752-3929).
     * Frame pictures only (MBAFF frames would take the field scan for field MBs). */
    int32_t  scaling_matrix;
    uint8_t  weight_scale4x4[6][16];          /* Intra Y, Cb, Cr, Inter Y, Cb, Cr      */
    uint8_t  weight_scale8x8[2][64];          /* Intra Y, Inter Y                      */
} h264g_pic_desc;

typedef struct h264g_ctx h264g_ctx;
typedef struct h264g_pic h264g_pic;

/* ------------------------------------------------------------- lifecycle */
/* Replaces: decoder construction + CH264PicturesGOP allocation
 * (H264VideoDecoder.cpp:56, H264PicturesGOP.cpp:37-47). */
int32_t h264g_create(const h264g_config *cfg, h264g_ctx **out);
int32_t h264g_destroy(h264g_ctx *ctx);

/* Reserve a stream's device DPB.  Fails with H264G_E_BUDGET when the HBM budget
 * would be exceeded (serving admission).  Replaces the per-frame my_malloc of
 * CH264PictureBase::init (H264PictureBase.cpp:160-182). */
int32_t h264g_stream_open(h264g_ctx *ctx, const h264g_stream_desc *desc, uint32_t *stream_id);
/* Returns the stream's device memory to the context (reused once the work enqueued so
 * far has completed; the call does not wait).  Pictures staged for the stream, and
 * batches holding them, can then no longer be launched: h264g_picture_submit,
 * h264g_batch_create / _submit and h264g_batch_run return H264G_E_STATE before any
 * launch, also when the id has been given to a newly opened stream.  Release and
 * destroy them as usual. */
int32_t h264g_stream_close(h264g_ctx *ctx, uint32_t stream_id);

/* Band heights (MB rows per band) of the BANDS schedules from now on, for the
 * autotuner; 0 keeps a height.  A stream's strips between bands are sized when it is
 * opened, so heights below those of any open stream fail with H264G_E_STATE, and a
 * batch created before the call fails in h264g_batch_run (H264G_E_STATE) until it is
 * created again.  No counterpart in the reference (it has no GPU schedules). */
int32_t h264g_set_bands(h264g_ctx *ctx, int32_t deblock_rows, int32_t intra_rows);

/* -------------------------------------------------------- plain path */
/* Stage one picture's syntax: copies the records and levels to device memory
 * (input arrays stay owned by the caller).  Replaces the per-MB dispatch at
 * H264SliceData.cpp:386-487 (the host calls this where the reference called the
 * reconstruction functions, once per picture).  Several threads may stage pictures
 * at once: only the stream lookup and the device block allocation are serialised. */
int32_t h264g_picture_begin(h264g_ctx *ctx, uint32_t stream_id, const h264g_pic_desc *desc,
                            const h264g_mb *mbs, uint32_t n_mbs,
                            const int16_t *levels, uint64_t n_levels, h264g_pic **out);

/* Enqueue reconstruction + deblocking of one picture on the context's HIP stream.
 * Replaces Deblocking_filter_process at H264PictureBase.cpp:707. */
int32_t h264g_picture_submit(h264g_ctx *ctx, h264g_pic *pic);

/* Allocation-free decode of one picture.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The picture's staged syntax and stage buffers
 * (records, levels, work lists, boundary strengths, intra residuals) go into a static
 * device array of H264G_STATIC_SCRATCH_BYTES that every device the library runs on holds
 * until the process ends, instead of the context's pooled device blocks; the launches and
 * the result are those of h264g_picture_begin + h264g_picture_submit.  The call makes no
 * hipMalloc / hipFree or other device allocation (the stream's DPB comes from
 * h264g_stream_open as usual).  A picture whose staging needs more than the array holds
 * fails with H264G_E_ARG before anything is copied (any 1920x1088 picture fits).  All
 * calls on one device share its array: each call first waits for the launches of the
 * previous call on that device, and concurrent calls are serialised, so a later call never
 * overwrites what an earlier one's kernels still read.  Returns once the launches are
 * enqueued (h264g_sync / fences wait for them). */
#define H264G_STATIC_SCRATCH_BYTES (64u << 20)
int32_t h264g_picture_decode_static(h264g_ctx *ctx, uint32_t stream_id, const h264g_pic_desc *desc,
                                    const h264g_mb *mbs, uint32_t n_mbs, const int16_t *levels, uint64_t n_levels);

/* Enqueue many staged pictures as one batched job: one persistent launch per
 * stage drains a device queue of macroblock tiles from all of them.  The
 * pictures must not depend on each other: pictures of different streams, and
 * up to H264G_MAX_STREAM_PICS pictures of one stream with distinct dst slots
 * none of which reads a slot another of them writes (e.g. the B pictures
 * between two P pictures together with the next P picture); H264G_E_ARG
 * otherwise.  No counterpart in the reference, which decodes one picture at a
 * time. */
#define H264G_MAX_STREAM_PICS 4
int32_t h264g_batch_submit(h264g_ctx *ctx, h264g_pic *const *pics, uint32_t n_pics);

/* The same in two steps: build the batch's device work queues once
 * (h264g_batch_create), then enqueue its stage launches (h264g_batch_run) as
 * often as wanted, e.g. inside a timed region with all inputs already in HBM. */
typedef struct h264g_batch h264g_batch;
int32_t h264g_batch_create(h264g_ctx *ctx, h264g_pic *const *pics, uint32_t n_pics, h264g_batch **out);
int32_t h264g_batch_run(h264g_ctx *ctx, h264g_batch *batch);
int32_t h264g_batch_destroy(h264g_ctx *ctx, h264g_batch *batch);

/* Device-side time per stage, in milliseconds, summed over every stage sequence
 * launched since the previous call (HIP events recorded around each stage on the
 * launch stream; a batch with pictures on both lanes launches one sequence per
 * lane, and the two lanes' sequences may overlap): out[0] inter, out[1] intra,
 * out[2] boundary strengths, out[3] deblocking.  Waits for both lanes; returns the
 * number of sequences summed.  Requires profiling (h264g_config.profile or
 * h264g_set_profile). */
int32_t h264g_stage_times(h264g_ctx *ctx, float out[4]);

/* Completion fences on the context's stream (serving loops): a fence records the
 * point after everything enqueued so far; query returns 1 when that work is done,
 * 0 when not yet, negative on error; wait blocks.  No counterpart in the reference,
 * which decodes synchronously. */
typedef struct h264g_fence h264g_fence;
int32_t h264g_fence_record(h264g_ctx *ctx, h264g_fence **out);
int32_t h264g_fence_query(h264g_ctx *ctx, h264g_fence *fence);
int32_t h264g_fence_wait(h264g_ctx *ctx, h264g_fence *fence);
int32_t h264g_fence_destroy(h264g_ctx *ctx, h264g_fence *fence);

/* Turn per-stage event recording on or off for later launches (as
 * h264g_config.profile).  While it is on, batches launch their stages directly
 * instead of replaying their captured graphs. */
int32_t h264g_set_profile(h264g_ctx *ctx, int32_t on);

/* Wait for all enqueued work; returns the first device error word if any. */
int32_t h264g_sync(h264g_ctx *ctx);

/* Copy a decoded DPB slot to host planes (stride = width).  Call before the
 * output callback fires, as do_callback does (H264VideoDecoder.cpp:383-435). */
int32_t h264g_slot_readback(h264g_ctx *ctx, uint32_t stream_id, int32_t slot,
                            uint8_t *y, uint8_t *cb, uint8_t *cr);
/* Convert a decoded DPB slot to BGR24 on the device and copy it to host memory:
 * rows of `stride` bytes (>= 3 * width), bottom-up when flip = 1.  Replaces
 * CH264PictureBase::convertYuv420pToBgr24 / ...FlipLines (H264PictureBase.cpp:440-495)
 * used by saveToBmpFile and the player. */
int32_t h264g_slot_to_bgr24(h264g_ctx *ctx, uint32_t stream_id, int32_t slot, uint8_t *bgr, uint32_t stride,
                            int32_t flip);

/* Upload host planes into a DPB slot (tests, resuming a stream at an IDR). */
int32_t h264g_slot_upload(h264g_ctx *ctx, uint32_t stream_id, int32_t slot,
                          const uint8_t *y, const uint8_t *cb, const uint8_t *cr);

/* Release a staged picture's device syntax buffers.  Does not wait: the memory is
 * reused only after the launches enqueued before the release have completed. */
int32_t h264g_picture_release(h264g_ctx *ctx, h264g_pic *pic);

/* Algorithmic bytes of a staged picture for the roofline report:
 * out[0] = MC + residual (inter), out[1] = intra, out[2] = deblocking. */
int32_t h264g_picture_alg_bytes(const h264g_pic *pic, uint64_t out[3]);

/* Measurement aid with no counterpart in the reference: runs the library's plain
 * copy kernel (16-byte loads and stores) `reps` times over a device buffer of
 * `bytes` (a multiple of 16) into another, and waits.  Its HBM counter readings
 * calibrate the FETCH_SIZE / WRITE_SIZE of the reconstruction kernels. */
int32_t h264g_calibration_copy(h264g_ctx *ctx, uint64_t bytes, int32_t reps);

/* Last error message of the calling thread (static storage). */
const char *h264g_last_error(void);

/* "<gcnArchName>|<device name>|<N> CUs|<clock> kHz" of a HIP device, for caches of
 * measured choices (the band-height autotuner).  No counterpart in the reference. */
int32_t h264g_device_describe(int32_t device, char *out, uint32_t len);

/* ----------------------------------------------------- syntax generator */
/* Deterministic synthetic syntax (configs of the benchmark).  Host-only. */
typedef struct h264g_gen h264g_gen;

typedef struct h264g_gen_params {
    int32_t  config;              /* 1 = CIF I+P, 2 = 1080p IPBBPBBP High, 3 = 1080p
                                     MBAFF (config 2 syntax, 4x4 transform only),
                                     0 = custom                                        */
    uint64_t seed;
    int32_t  width_mbs, height_mbs;
    int32_t  n_frames;
    int32_t  gop;                 /* frames between IDRs (0 = one GOP)                */
    int32_t  b_frames;            /* B pictures between anchors                        */
    int32_t  max_refs;
    int32_t  t8x8;                /* allow transform_size_8x8_flag / Intra8x8          */
    int32_t  mv_mode;             /* 0 uniform [-mv_range, mv_range], 1 Gaussian       */
    int32_t  mv_range;            /* uniform half-range, or Gaussian sigma (qpel)      */
    int32_t  mv_clamp_lo, mv_clamp_hi;
    int32_t  qp_lo, qp_hi;
    int32_t  zero_permille;       /* probability a level is zero, in 1/1000            */
    int32_t  laplace_b_x100;      /* Laplacian scale of level magnitudes, x100         */
    int32_t  intra_in_inter_permille;
    int32_t  skip_permille;
    int32_t  weighted_pred_flag, weighted_bipred_idc;
    int32_t  filter_offsets;      /* draw slice alpha/beta offsets in [-2,2]           */
    int32_t  disable_deblocking_filter_idc;
    int32_t  chroma_qp_index_offset, second_chroma_qp_index_offset;
    int32_t  ipcm_permille;
    int32_t  p_pattern;           /* 1 = config-1 P mix (60/30/10), 0 = general mix    */
    int32_t  spare_slots;         /* DPB slots beyond max_refs + 1; with any, pictures
                                     take the free slot after the last one used instead
                                     of the lowest, so a picture does not overwrite the
                                     slot of the picture decoded just before it        */
    int32_t  mbaff;               /* 1 = MBAFF frames: MB pairs, each pair a field pair
                                     with probability 1/2 (the inferred flag when both
                                     MBs are skipped); height_mbs must be even          */
    int32_t  constrained_intra_pred; /* constrained_intra_pred_flag: intra MBs predict
                                     only from intra neighbours (frames only)           */
    int32_t  direct_permille;     /* B pictures: share of spatial-direct MBs (B_Skip,
                                     B_Direct_16x16, B_8x8 with B_Direct_8x8 sub-MBs in
                                     equal parts; frames, no implicit weights)          */
    int32_t  slices;              /* slices per picture (0 or 1: one), runs of
                                     ceil(W * H / slices) MBs in address order; intra
                                     prediction sees no neighbour of another slice.
                                     Frames only (no MBAFF)                              */
} h264g_gen_params;

/* ------------------------------------------------ host motion-vector derivation
 * The final motion vectors of an h264g_mb record from the parsed syntax, for frame
 * pictures: P_Skip vectors (8.4.1.1) and mvp + mvd with the luma motion-vector
 * prediction over the neighbouring partitions (8.4.1.3, 6.4.11.7).  Replaces what the
 * reference derives inside reconstruction: P_Skip at H264InterPrediction.cpp:685-738,
 * Derivation_process_for_luma_motion_vector_prediction and its neighbouring
 * partitions at H264InterPrediction.cpp:1660-2001; spatial direct prediction after
 * h264g_mvfield_colocated.  Host only, no device work; one field per picture being
 * parsed, not thread-safe. */
typedef struct h264g_mvfield h264g_mvfield;
int32_t h264g_mvfield_create(int32_t width_mbs, int32_t height_mbs, h264g_mvfield **out);
void    h264g_mvfield_destroy(h264g_mvfield *field);
/* Start of a picture: no MB decoded yet. */
int32_t h264g_mvfield_reset(h264g_mvfield *field);
/* MB mb_addr of the picture, called for every MB in decoding (raster) order, intra
 * MBs included.  In: mb_type, part_shape, sub_part_shape, ref_idx (P_Skip: set to
 * refIdxL0 = 0 here), slice_num (neighbours of another slice are not available).  mvd: int16 [list][16][2], the mvd of each (sub-)partition and
 * list at its top-left 4x4 block (raster index, the layout of mb->mv), read for the
 * lists with ref_idx >= 0; NULL for intra and P_Skip MBs.  Out: mb->mv of every 4x4
 * block of the used lists.  H264G_E_ARG for a type the field does not derive. */
int32_t h264g_mvfield_derive(h264g_mvfield *field, uint32_t mb_addr, h264g_mb *mb, const int16_t *mvd);
/* Intra4x4PredMode / Intra8x8PredMode of I_NxN MB mb_addr from its parsed syntax
 * (8.3.1.1 / 8.3.2.1; replaces getIntra4x4PredMode / getIntra8x8PredMode,
 * H264PictureBase.cpp:773-1058): predIntraNxNPredMode = Min(mode of neighbour A, of B),
 * DC (2) when a neighbour is not available, is inter under constrained intra prediction
 * or is not I_NxN; an I4x4 neighbour of an 8x8 block gives its 4x4 block 1 (left) or 2
 * (above) of the neighbouring 8x8 block.  Call it before h264g_mvfield_derive of the same
 * MB (which records the modes for the MBs after it).  In: mb_type (I4x4 or I8x8);
 * syntax: uint8 [16][2] (prev_intra*_pred_mode_flag, rem_intra*_pred_mode) per
 * luma4x4BlkIdx, or per luma8x8BlkIdx in the first four entries for I8x8;
 * constrained_intra_pred: the PPS flag.  Out: mb->intra_pred_modes. */
int32_t h264g_mvfield_intra_modes(h264g_mvfield *field, uint32_t mb_addr, h264g_mb *mb, const uint8_t *syntax,
                                  int32_t constrained_intra_pred);
/* The co-located picture of the B picture being parsed, for spatial direct prediction
 * (8.4.1.2.1 / 8.4.1.2.2; replaces the reference's colocated lookup and spatial direct
 * derivation, H264InterPrediction.cpp:922-1432): col_mbs are the final records of
 * RefPicList1[0] (W x H, raster), col_short_term whether it is marked "used for
 * short-term reference"; NULL clears it.  Frame pictures, direct_8x8_inference_flag 1,
 * direct_spatial_mv_pred_flag 1 only (the reference's temporal direct divides by zero).
 * Afterwards h264g_mvfield_derive fills ref_idx and mv of the quadrants of H264G_MB_B
 * records that carry H264G_MBF_DIRECT bits (mvd may then be NULL when all four do). */
int32_t h264g_mvfield_colocated(h264g_mvfield *field, const h264g_mb *col_mbs, int32_t col_short_term);

/* ------------------------------------------------------------ bitstream front end
 * Annex B byte stream of frame pictures coded with CAVLC or CABAC, in one slice or several
 * per picture: NAL units, SPS / PPS, slice headers, slice data and the decoded picture
 * buffer (POC types 0, 1 and 2, sliding-window and adaptive reference marking, reference
 * list initialisation and modification, DPB slots).  Replaces the reference's parsing and
 * picture management on the way to reconstruction: CH264NalUnit / CH264SPS / CH264PPS /
 * CH264SliceHeader (H264NalUnit.cpp:122, H264SliceHeader.cpp:289-509),
 * CH264SliceData::slice_data (H264SliceData.cpp:386-487), CH264MacroBlock::macroblock_layer
 * (H264MacroBlock.cpp:912-1139), CH264ResidualBlockCavlc (H264ResidualBlockCavlc.cpp),
 * CH264Cabac (H264Cabac.cpp) and H264RefPicList.cpp.
 * Slices of one picture must share one descriptor (slice type, reference lists, weights,
 * deblocking parameters).  Slice data that fails to parse ends its slice where it fails, and MBs that no
 * slice decodes are H264G_MBF_LOST records, as the reference skips the rest of such a
 * slice and decodes on (H264VideoDecoder.cpp:135-136); lost_mbs counts them.
 * SPS / PPS scaling matrices give the descriptor's lists (h264g_pic_desc.scaling_matrix).
 * Refused with H264G_E_ARG: field pictures and MBAFF, slice groups, data partitioning,
 * formats other than 8-bit 4:2:0, lossless coding, SP / SI slices, temporal direct
 * prediction.  A gap in frame_num inserts no frames (the reference's gap process is
 * empty).  Host only. */
typedef struct h264g_parser h264g_parser;
typedef struct h264g_parsed_picture {
    h264g_pic_desc desc;          /* ready for h264g_picture_begin (dst_slot, ref_slot ...) */
    const h264g_mb *mbs;          /* final records (motion vectors and intra modes derived) */
    uint32_t n_mbs;
    const int16_t *levels;
    uint64_t n_levels;            /* the three stay valid until the next h264g_parser_next */
    int32_t width_mbs, height_mbs;
    int32_t dpb_slots;            /* slots the stream needs: max_num_ref_frames + 1        */
    int32_t max_num_reorder_frames;  /* from the VUI, -1 when absent                       */
    int32_t idr, nal_ref_idc, frame_num;
    int32_t profile_idc, constraint_set3, level_idc;
    int32_t n_slices;             /* slices that made the picture                          */
    int32_t lost_mbs;             /* H264G_MBF_LOST records (0 for an intact picture)       */
    int32_t first_lost_mb;        /* address of the first of them, -1 when none             */
} h264g_parsed_picture;
/* The parser keeps a copy of the byte stream.  Pictures after the one handed out are
 * gathered ahead and their slice data parsed on worker threads, one per picture: as many
 * as the environment variable H264G_PARSE_THREADS names when the parser is created (1:
 * every picture parses on the calling thread), else the host's hardware threads, at most
 * 8.  A worker derives the vectors and modes of a picture without B slices; those of B
 * pictures (co-located direct prediction) are derived on the calling thread in decoding
 * order.  The pictures, errors and damage reports are those of the in-order parse. */
int32_t h264g_parser_create(const uint8_t *data, uint64_t len, h264g_parser **out);
int32_t h264g_parser_destroy(h264g_parser *parser);
/* The next picture in decoding order: 1 with *out filled, 0 at the end of the stream,
 * negative on an error (h264g_last_error names the picture and MB). */
int32_t h264g_parser_next(h264g_parser *parser, h264g_parsed_picture *out);
/* Why the last picture's damaged slice stopped (the parse error at its first_lost_mb),
 * or "" when its slices were intact; valid until the next h264g_parser_next. */
const char *h264g_parser_damage(const h264g_parser *parser);

int32_t h264g_gen_default_params(int32_t config, h264g_gen_params *out);
int32_t h264g_gen_create(const h264g_gen_params *params, h264g_gen **out);
int32_t h264g_gen_destroy(h264g_gen *gen);
/* Pictures are produced in decoding order. */
int32_t h264g_gen_num_pictures(const h264g_gen *gen);
int32_t h264g_gen_picture(h264g_gen *gen, int32_t idx, h264g_pic_desc *desc,
                          const h264g_mb **mbs, uint32_t *n_mbs,
                          const int16_t **levels, uint64_t *n_levels);
/* Number of DPB slots the stream needs. */
int32_t h264g_gen_dpb_slots(const h264g_gen *gen);
/* FNV-1a-64 of all generated syntax (determinism check across machines). */
uint64_t h264g_gen_digest(const h264g_gen *gen);

#ifdef __cplusplus
}
#endif

#endif /* H264G_H */
