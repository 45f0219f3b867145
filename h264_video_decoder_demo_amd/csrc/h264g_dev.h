// h264g_dev.h -- layouts shared by the host side of libh264g and its gfx950 kernels.
#pragma once
#include <stdint.h>
#include "h264g.h"

// Per-MB values the host derives once while staging a picture (picture_begin):
// everything the kernels would otherwise recompute per sample.
struct DevAux {
    int8_t   refid[2][4];     // reference picture per list and 8x8 quadrant as a DPB slot,
                              // -1 = none: the MC source when the list is used, and (frame
                              // pictures) the picture identity the bS derivation compares;
                              // field MBs of MBAFF frames: the frame of refIdx / 2
    uint8_t  pf;              // PredFlagLX of quadrant q at bit (4 * X + q)
    uint8_t  kind;            // AUX_* bits below
    uint16_t nz4;             // luma4x4BlkIdx b has non-zero levels (4x4-transform MBs)
    uint8_t  wp[4];           // weight-table entry per quadrant
    uint8_t  nz8;             // 8x8 block has non-zero levels (8x8-transform MBs)
    uint8_t  nmask;           // number of level masks of the MB
    uint8_t  qpc[2];          // QP'C of Cb / Cr for dequantisation
    uint8_t  qdb[3];          // deblocking qP: luma, Cb edges, Cr edges
    uint8_t  fpar;            // field MBs of MBAFF frames: bit (4 * X + q) = the referenced
                              // field is the bottom field (refIdx odd flips the MB's parity)
    uint32_t ires;            // bits 0..21: intra MBs, index of the MB's residual in
                              // DevPic::ires; bits 22..31: number of non-zero level values
};
#define AUX_IRES_MASK    0x3FFFFFu
#define AUX_NVALS_SHIFT  22
static_assert(sizeof(DevAux) == 28, "DevAux layout");

#define AUX_INTRA   0x01u  // Intra_4x4 / Intra_8x8 / Intra_16x16 prediction (bS 3/4)
#define AUX_PCM     0x02u
#define AUX_T8      0x04u  // transform_size_8x8_flag
#define AUX_INTER   0x08u
#define AUX_RESID   0x10u  // has coefficient levels
// Entry of an intra MB in the intra band schedule (DevPic::irow after the row offsets):
// column in bits 0..10, neighbour availability (A, B, C, D) from ICOL_AV_SHIFT, the
// available neighbours that are intra MBs the MB waits for from ICOL_DEP_SHIFT, and
// ICOL_RES when the MB has levels (its residual is at the entry's index in DevPic::ires).
#define ICOL_X_MASK    0x7FFu
#define ICOL_AV_SHIFT  11
#define ICOL_DEP_SHIFT 15
#define ICOL_RES       (1u << 19)

// Resolved weighted-prediction parameters of one (refIdxL0, refIdxL1) combination.
struct DevWeight {
    int16_t mode;             // 0 default (8.4.2.3.1), 1 weighted (8.4.2.3.2)
    int16_t logWD[3];
    int16_t w0[3], w1[3], o0[3], o1[3];
    // L1-only luma takes (w0, o0) -- handled by the host filling w0/o0 for such entries
};

// One picture as the kernels see it.
struct DevPic {
    const h264g_mb *mbs;
    const DevAux *aux;
    const int16_t *levels;
    const DevWeight *wt;
    uint8_t *dpb;             // slot 0 of the stream's DPB
    uint64_t slot_bytes;      // bytes per DPB slot (Y + Cb + Cr)
    uint32_t *flags;          // per-MB wavefront progress words of the stream
    uint64_t *rowprog;        // per-MB-row progress of row-pipelined deblocking:
                              // epoch << 32 | MBs of the row finished
    uint8_t  *bs;             // boundary strengths, 32 per MB: [dir][4-line segment][edge]
    const uint32_t *irow;     // intra MBs per row: H + 1 offsets, then one ICOL entry per intra MB
    uint32_t n_intra;         // intra-predicted MBs of the picture
    int32_t  dst;             // destination slot
    int32_t  W, H;            // in macroblocks
    int32_t  offA, offB;      // FilterOffsetA / FilterOffsetB
    int32_t  dbk_idc;         // disable_deblocking_filter_idc
    int32_t  cip;             // constrained_intra_pred_flag
    uint8_t  *gstrip;         // deblocking strips between bands of the stream: W x 256 B per band
    uint64_t *istrip;         // intra strips between bands of the stream: W x 8 words per band
    uint32_t *dbq;            // per-MB deblocking word, written by k_bs
    int16_t  *ires;           // residual of each intra MB with levels (DevAux::ires): luma
                              // 16x16 then Cb, Cr 8x8, raster; written by the inter stage
    int32_t  mbaff;           // MbaffFrameFlag: MB address 2 p + b is the top / bottom MB of
                              // pair p, pairs in raster order
    const int8_t *bsref;      // MBAFF frames: per MB [list][quadrant] the picture identity
                              // the bS derivation compares, RefPicListX[RefIdxLX] (-1: none)
    uint32_t *pairprog;       // MBAFF streams: per-MB-pair deblocking progress words
    uint32_t dbk_stop;        // MBs from this address on filter none of their edges (the
                              // first MB of no slice; W * H when every MB has one)
    int32_t  multi_slice;     // the MBs carry more than one slice number (MBs of no slice
                              // aside): k_bs reads them
    const uint16_t *ls;       // scaling matrices: LevelScale4x4 [intra, inter luma list][qP % 6]
                              // [raster 16] at 0, LevelScale8x8 [list][qP % 6][raster 64] at
                              // LS8_AT; null with flat lists
};
#define LS4_LIST 96           // entries per 4x4 list (6 x 16)
#define LS8_AT   192          // first 8x8 entry
#define LS8_LIST 384          // entries per 8x8 list (6 x 64)
#define LS_ENTRIES 960

// Work item: picture index in the batch and macroblock address.
#define WI_PIC_SHIFT 20u
#define WI_MB_MASK   ((1u << WI_PIC_SHIFT) - 1u)

// Control block of one persistent launch: ticket dispenser + finished-wave count.
struct DevCtl {
    uint32_t ticket;
    uint32_t done;
    uint32_t pad[14];         // own 64-byte line
};

// Device error bits (ORed into the context's error word).
#define DEVERR_TIMEOUT  0x1u

extern "C" {
// Kernel launchers (kernels.hip); all enqueue on `stream` and return a hipError_t.
// mbaff: every picture of the launch is an MBAFF frame (else none is).
// scaled: some picture of the launch has scaling matrices (DevPic::ls).
int h264g_launch_inter(const DevPic *pics, const uint32_t *items, uint32_t n, const uint32_t *intra_items, uint32_t n_intra,
                       int mbaff, int scaled, uint32_t *err, void *stream);
// `epoch` points at the device word shared by all wavefront launches of a context
// (initially 1; each launch's last wave advances it).
int h264g_launch_intra(const DevPic *pics, const uint32_t *items, uint32_t n, DevCtl *ctl, uint32_t *epoch,
                       uint32_t *err, int waves, int mbaff, void *stream);
int h264g_launch_deblock(const DevPic *pics, const uint32_t *items, uint32_t n, DevCtl *ctl, uint32_t *epoch,
                         uint32_t *err, int waves, void *stream);
int h264g_launch_deblock_rows(const DevPic *pics, const uint32_t *items, uint32_t n, DevCtl *ctl, uint32_t *epoch,
                              uint32_t *err, int waves, void *stream);
// One workgroup of `waves` waves per picture (pics[0..n_pics)); pictures of up to 1024 MB rows.
int h264g_launch_intra_band(const DevPic *pics, const uint32_t *bands, uint32_t n, DevCtl *ctl, uint32_t *epoch,
                            int waves, uint32_t *err, void *stream);
// One workgroup of ceil(rows / 2) waves per band of `rows` MB rows (two rows per wave).
int h264g_launch_deblock_band(const DevPic *pics, const uint32_t *bands, uint32_t n, DevCtl *ctl, uint32_t *epoch,
                              int rows, uint32_t *err, void *stream);
int h264g_launch_bs(const DevPic *pics, const uint32_t *items, uint32_t n, void *stream);
// Deblocking of MBAFF frames: persistent waves over MB pairs (items: picture index << 20 |
// pair index, ordered by t = x + 2 y of the pair), one pair per wave at a time.
int h264g_launch_deblock_mbaff(const DevPic *pics, const uint32_t *items, uint32_t n, DevCtl *ctl, uint32_t *epoch,
                               uint32_t *err, int waves, void *stream);
int h264g_launch_bgr24(const uint8_t *yuv, int wl, int hl, uint8_t *bgr, uint32_t stride, int flip, void *stream);
int h264g_launch_copy(const uint8_t *src, uint8_t *dst, uint64_t bytes, void *stream);
}
// Device address of the static scratch array (H264G_STATIC_SCRATCH_BYTES) on `device`,
// which must be the current device; looked up on the first call per device.  Returns a
// hipError_t.
int h264g_scratch_address(int device, void **out);
