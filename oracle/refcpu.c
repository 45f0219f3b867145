/*
 * refcpu.c -- TEST INFRASTRUCTURE ONLY: the oracle.
 *
 * Plain-C restatement of the picture reconstruction of jfu222/h264_video_decoder_demo
 * for progressive frames (one slice or several per picture): dequantisation + inverse
 * transforms (8.5), inter prediction with fractional-sample interpolation and
 * weighted prediction (8.4.2 / 8.4.3), intra prediction (8.3) and the deblocking
 * filter (8.7), including the reference's deterministic deviations from the spec:
 *   Q1  chroma DC prediction treats a neighbour sample equal to 0 as unavailable
 *       (H264PictureBase.cpp:2205-2256);
 *   Q2  explicit WP, L1-only partitions: luma uses (w, o) = (luma_weight_l0_flag,
 *       luma_weight_l0[31]) read through index -1 (H264InterPrediction.cpp:2764,2768,
 *       3011-3013 with the member layout of H264SliceHeader.h:78-80);
 *   Q3  bS = 1 test for bi-predicted partitions with one reference picture parses as
 *       L0L0 || (L1L1 && L0pL1q) || L1pL0q (Deblock.cpp:1289-1299);
 *   Q13 with scaling matrices, chroma blocks dequantise with the LevelScale4x4 of the
 *       MB's luma list (Intra Y or Inter Y), the one its luma driver derived last: the
 *       chroma drivers never derive their own (H264PictureBase.cpp:3752-3929, 4852-4989);
 *   Q16 deblocking derives the Cb edge QPs with second_chroma_qp_index_offset and the
 *       Cr edge QPs with chroma_qp_index_offset (argument mix-up, Deblock.cpp:813-815);
 *   I_PCM macroblocks are not "intra" for bS (m_mb_pred_mode is Intra_NA, so
 *       IS_INTRA_Prediction_Mode is false, Deblock.cpp:1022-1102; H264CommonFunc.h:272)
 *       and take RefIdxLX = 0 / PredFlagLX = 0 from the zeroed MB state.
 * MBAFF frames (macroblock pairs, 6.4.12.2 / 8.7 with field macroblocks) follow the
 * reference too, with its deviations there:
 *   Q4  bS compares RefPicListX[RefIdxLX] of field MBs, whose refIdx addresses the
 *       field list (Deblock.cpp:1175-1178), so the entry is refIdx, not refIdx / 2;
 *   Q5  frame MB under a field pair: the first chroma top-edge pass takes its p-MB
 *       data from the bottom field MB where luma takes the top one (Deblock.cpp:474-481
 *       vs 264);
 *   Q6  left edge of a frame MB beside a field pair: the p-MB alternates per line, but
 *       the p block for bS 2 uses the frame line number (Deblock.cpp:755, 773); beside a
 *       frame pair a field MB takes every line's p data from the top frame MB;
 *   Q17 field-MB interpolation clamps x to the field view's row pitch, twice the width
 *       (H264InterPrediction.cpp:2363, 2504-2507; H264PictureBase.cpp:209-210): samples
 *       right of the picture come from the next row of the frame buffer, on the bottom
 *       field's last row from the next plane or past the Cr plane (the caller's planes
 *       carry a zeroed tail, see refcpu.h).
 *   Q18 constrained intra prediction hides an inter MB only when its m_mb_pred_mode is
 *       Pred_L0 / Pred_L1 / BiPred (IS_INTER_Prediction_Mode, H264CommonFunc.h:273;
 *       H264PictureBase.cpp:1129, 1492, 1905, 2158): P_8x8 / B_8x8 MBs (Pred_NA,
 *       H264MacroBlock.cpp:99-100, 142) stay available to intra prediction.
 * Slices: a neighbour MB of another slice (h264g_mb.slice_num) is not available
 * (H264PictureBase.cpp:2890-2937); with disable_deblocking_filter_idc 2 MB edges
 * between slices are not filtered (Deblock.cpp:128-153);
 *   Q7  with disable_deblocking_filter_idc 0 such an edge is filtered with the current MB
 *       as its p-MB (the unavailable neighbour leaves the p-MB address at the current one,
 *       Deblock.cpp:747-760, 784-791): bS of an internal edge between the MB's blocks at
 *       x (or y) 15 and 0, the MB's own qP on both sides, the neighbour's samples.
 *   Q22 the deblocking of a picture stops at its first MB of no slice
 *       (H264G_SLICE_NONE, an MB that no slice decoded): the bS derivation fails on the
 *       MB's reset type MB_TYPE_NA (H264InterPrediction.cpp:3182-3186) and
 *       end_decode_the_picture ignores the failure (H264PictureBase.cpp:707-708), so
 *       that MB and every later one filter none of their edges.
 * Each function cites the reference code it follows.  Used only by tests/, smoke()
 * and bench.py's CPU baseline, as the checker.
 */
#include "refcpu.h"

#include <stdlib.h>
#include <string.h>

#define CLIP3(lo, hi, v) ((v) < (lo) ? (lo) : ((v) > (hi) ? (hi) : (v)))
#define CLIP1(v) CLIP3(0, 255, (v))
static int iabs(int v) { return v < 0 ? -v : v; }
static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------ tables */
/* Table 8-13: scan position -> raster index (row * 4 + col), zig-zag and field scan
 * (H264PictureBase.cpp:4542-4593). */
static const uint8_t ZZ4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
static const uint8_t FLD4[16] = {0, 4, 1, 8, 12, 5, 9, 13, 2, 6, 10, 14, 3, 7, 11, 15};
/* Table 8-14 (H264PictureBase.cpp:4597-4744). */
static const uint8_t ZZ8[64] = {
    0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14, 21, 28,
    35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
static const uint8_t FLD8[64] = {
    0, 8, 16, 1, 9, 24, 32, 17, 2, 25, 40, 48, 56, 33, 10, 3, 18, 41, 49, 57, 26, 11, 4, 19, 34, 42, 50, 58, 27, 12, 5, 20,
    35, 43, 51, 59, 28, 13, 6, 21, 36, 44, 52, 60, 29, 14, 22, 37, 45, 53, 61, 30, 7, 15, 38, 46, 54, 62, 23, 31, 39, 47, 55, 63};
/* normAdjust4x4 / normAdjust8x8 (8-315, 8-318; H264PictureBase.cpp:4900-4950) */
static const int V4[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16}, {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
static const int V8[6][6] = {{20, 18, 32, 19, 25, 24}, {22, 19, 35, 21, 28, 26}, {26, 23, 42, 24, 33, 31},
                             {28, 25, 45, 26, 35, 33}, {32, 28, 51, 30, 40, 38}, {36, 32, 58, 34, 46, 43}};
/* Table 8-15 (H264PictureBase.cpp:4773-4774) */
static const uint8_t QPC_TAB[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};
/* Table 8-16 (Deblock.cpp:1330-1348) */
static const uint8_t ALPHA[52] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 4, 4, 5, 6, 7, 8, 9, 10, 12, 13,
                                  15, 17, 20, 22, 25, 28, 32, 36, 40, 45, 50, 56, 63, 71, 80, 90, 101, 113, 127, 144, 162, 182, 203, 226, 255, 255};
static const uint8_t BETA[52] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4,
                                 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15, 15, 16, 16, 17, 17, 18, 18};
/* Table 8-17 (Deblock.cpp:1383-1409) */
static const uint8_t TC0[3][52] = {
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 6, 6, 7, 8, 9, 10, 11, 13},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5, 5, 6, 7, 8, 8, 10, 11, 12, 13, 15, 17},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 23, 25}};

/* LevelScale4x4 / 8x8 (8-315 / 8-318, H264PictureBase.cpp:4852-4989): weightScale (raster,
 * NULL = the flat 16) times normAdjust */
static int ls4w(const uint8_t *w, int m, int i, int j) {
    int cls = (i % 2 == 0 && j % 2 == 0) ? 0 : ((i % 2 == 1 && j % 2 == 1) ? 1 : 2);
    return (w ? w[i * 4 + j] : 16) * V4[m][cls];
}
static int ls4(int m, int i, int j) { return ls4w(NULL, m, i, j); }
static int ls8w(const uint8_t *w, int m, int i, int j) {
    int cls;
    if (i % 4 == 0 && j % 4 == 0) cls = 0;
    else if (i % 2 == 1 && j % 2 == 1) cls = 1;
    else if (i % 4 == 2 && j % 4 == 2) cls = 2;
    else if ((i % 4 == 0 && j % 2 == 1) || (i % 2 == 1 && j % 4 == 0)) cls = 3;
    else if ((i % 4 == 0 && j % 4 == 2) || (i % 4 == 2 && j % 4 == 0)) cls = 4;
    else cls = 5;
    return (w ? w[i * 8 + j] : 16) * V8[m][cls];
}

/* ------------------------------------------------- 8.5 transforms */
int refcpu_chroma_qp(int qpy, int offset) { /* H264PictureBase.cpp:4748-4813 */
    int qpi = CLIP3(0, 51, qpy + offset);
    return qpi < 30 ? qpi : QPC_TAB[qpi - 30];
}

/* 8.5.12 (H264PictureBase.cpp:4106-4266); r is raster row-major; w: weightScale4x4 or NULL. */
static void residual4x4_w(const int32_t scan[16], int qp, int dc_passthrough, int field, const uint8_t *w,
                          int32_t r[16]) {
    int32_t c[16], d[16], f[16];
    const uint8_t *tab = field ? FLD4 : ZZ4;
    for (int k = 0; k < 16; k++) c[tab[k]] = scan[k];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            int32_t v = c[i * 4 + j];
            if (i == 0 && j == 0 && dc_passthrough) d[0] = v;
            else if (qp >= 24) d[i * 4 + j] = (v * ls4w(w, qp % 6, i, j)) << (qp / 6 - 4);
            else d[i * 4 + j] = (v * ls4w(w, qp % 6, i, j) + (1 << (3 - qp / 6))) >> (4 - qp / 6);
        }
    for (int i = 0; i < 4; i++) { /* rows */
        int32_t *q = &d[i * 4];
        int32_t e0 = q[0] + q[2], e1 = q[0] - q[2], e2 = (q[1] >> 1) - q[3], e3 = q[1] + (q[3] >> 1);
        f[i * 4 + 0] = e0 + e3; f[i * 4 + 1] = e1 + e2; f[i * 4 + 2] = e1 - e2; f[i * 4 + 3] = e0 - e3;
    }
    for (int j = 0; j < 4; j++) { /* columns */
        int32_t g0 = f[j] + f[8 + j], g1 = f[j] - f[8 + j], g2 = (f[4 + j] >> 1) - f[12 + j], g3 = f[4 + j] + (f[12 + j] >> 1);
        r[j] = (g0 + g3 + 32) >> 6; r[4 + j] = (g1 + g2 + 32) >> 6;
        r[8 + j] = (g1 - g2 + 32) >> 6; r[12 + j] = (g0 - g3 + 32) >> 6;
    }
}
void refcpu_residual4x4(const int32_t scan[16], int qp, int dc_passthrough, int field, int32_t r[16]) {
    residual4x4_w(scan, qp, dc_passthrough, field, NULL, r);
}

static void idct8_1d(const int32_t *in, int stride, int32_t *out, int ostride) { /* 8.5.13.2 */
    int32_t d0 = in[0], d1 = in[stride], d2 = in[2 * stride], d3 = in[3 * stride];
    int32_t d4 = in[4 * stride], d5 = in[5 * stride], d6 = in[6 * stride], d7 = in[7 * stride];
    int32_t e0 = d0 + d4, e1 = -d3 + d5 - d7 - (d7 >> 1), e2 = d0 - d4, e3 = d1 + d7 - d3 - (d3 >> 1);
    int32_t e4 = (d2 >> 1) - d6, e5 = -d1 + d7 + d5 + (d5 >> 1), e6 = d2 + (d6 >> 1), e7 = d3 + d5 + d1 + (d1 >> 1);
    int32_t f0 = e0 + e6, f1 = e1 + (e7 >> 2), f2 = e2 + e4, f3 = e3 + (e5 >> 2);
    int32_t f4 = e2 - e4, f5 = (e3 >> 2) - e5, f6 = e0 - e6, f7 = e7 - (e1 >> 2);
    out[0] = f0 + f7; out[ostride] = f2 + f5; out[2 * ostride] = f4 + f3; out[3 * ostride] = f6 + f1;
    out[4 * ostride] = f6 - f1; out[5 * ostride] = f4 - f3; out[6 * ostride] = f2 - f5; out[7 * ostride] = f0 - f7;
}

/* 8.5.13 (H264PictureBase.cpp:4270-4403); w: weightScale8x8 or NULL */
static void residual8x8_w(const int32_t scan[64], int qp, int field, const uint8_t *w, int32_t r[64]) {
    int32_t c[64], d[64], g[64], m[64];
    const uint8_t *tab = field ? FLD8 : ZZ8;
    for (int k = 0; k < 64; k++) c[tab[k]] = scan[k];
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            int32_t v = c[i * 8 + j];
            if (qp >= 36) d[i * 8 + j] = (v * ls8w(w, qp % 6, i, j)) << (qp / 6 - 6);
            else d[i * 8 + j] = (v * ls8w(w, qp % 6, i, j) + (1 << (5 - qp / 6))) >> (6 - qp / 6);
        }
    for (int i = 0; i < 8; i++) idct8_1d(&d[i * 8], 1, &g[i * 8], 1);
    for (int j = 0; j < 8; j++) idct8_1d(&g[j], 8, &m[j], 8);
    for (int k = 0; k < 64; k++) r[k] = (m[k] + 32) >> 6;
}
void refcpu_residual8x8(const int32_t scan[64], int qp, int field, int32_t r[64]) { residual8x8_w(scan, qp, field, NULL, r); }

/* 8.5.10 (H264PictureBase.cpp:4993-5088; the DC levels take the MB's scan,
 * H264PictureBase.cpp:3519) */
static void luma_dc_scan(const int32_t scan[16], int qp, int field, const uint8_t *w, int32_t dcY[16]) {
    int32_t c[16], g[16], f[16];
    const uint8_t *tab = field ? FLD4 : ZZ4;
    for (int k = 0; k < 16; k++) c[tab[k]] = scan[k];
    for (int j = 0; j < 4; j++) { /* combine rows of c (first index) */
        int32_t a = c[j], b = c[4 + j], cc = c[8 + j], d = c[12 + j];
        g[j] = a + b + cc + d; g[4 + j] = a + b - cc - d; g[8 + j] = a - b - cc + d; g[12 + j] = a - b + cc - d;
    }
    for (int i = 0; i < 4; i++) {
        int32_t a = g[i * 4], b = g[i * 4 + 1], cc = g[i * 4 + 2], d = g[i * 4 + 3];
        f[i * 4] = a + b + cc + d; f[i * 4 + 1] = a + b - cc - d; f[i * 4 + 2] = a - b - cc + d; f[i * 4 + 3] = a - b + cc - d;
    }
    int s = ls4w(w, qp % 6, 0, 0);
    for (int k = 0; k < 16; k++) {
        if (qp >= 36) dcY[k] = (f[k] * s) << (qp / 6 - 6);
        else dcY[k] = (f[k] * s + (1 << (5 - qp / 6))) >> (6 - qp / 6);
    }
}
void refcpu_luma_dc(const int32_t scan[16], int qp, int32_t dcY[16]) { luma_dc_scan(scan, qp, 0, NULL, dcY); }

/* 8.5.11 for 4:2:0 (H264PictureBase.cpp:3989-4043) */
static void chroma_dc_w(const int32_t c[4], int qpc, const uint8_t *w, int32_t dcC[4]) {
    int32_t e00 = c[0] + c[2], e01 = c[1] + c[3], e10 = c[0] - c[2], e11 = c[1] - c[3];
    int32_t f[4] = {e00 + e01, e00 - e01, e10 + e11, e10 - e11};
    for (int k = 0; k < 4; k++) dcC[k] = ((f[k] * ls4w(w, qpc % 6, 0, 0)) << (qpc / 6)) >> 5;
}
void refcpu_chroma_dc(const int32_t c[4], int qpc, int32_t dcC[4]) { chroma_dc_w(c, qpc, NULL, dcC); }

/* ---------------------------------------------- 8.4.2.2 interpolation */
static int px(const uint8_t *ref, int w, int h, int x, int y) {
    return ref[CLIP3(0, h - 1, y) * w + CLIP3(0, w - 1, x)];
}
static int tap6(int a, int b, int c, int d, int e, int f) { return a - 5 * b + 20 * c + 20 * d - 5 * e + f; }

/* 8.4.2.2.1 (H264InterPrediction.cpp:2344-2480): per-sample, coordinates clamped. */
int refcpu_luma_sample(const uint8_t *ref, int w, int ht, int xI, int yI, int xF, int yF) {
#define S(dx, dy) px(ref, w, ht, xI + (dx), yI + (dy))
    int G = S(0, 0), H = S(1, 0), M = S(0, 1);
    if (xF == 0 && yF == 0) return G;
    /* full-sample-distance intermediates */
    int b1 = tap6(S(-2, 0), S(-1, 0), G, H, S(2, 0), S(3, 0));
    int h1 = tap6(S(0, -2), S(0, -1), G, M, S(0, 2), S(0, 3));
    int s1 = tap6(S(-2, 1), S(-1, 1), M, S(1, 1), S(2, 1), S(3, 1));
    int m1 = tap6(S(1, -2), S(1, -1), H, S(1, 1), S(1, 2), S(1, 3));
    int b = CLIP1((b1 + 16) >> 5), h = CLIP1((h1 + 16) >> 5), s = CLIP1((s1 + 16) >> 5), m = CLIP1((m1 + 16) >> 5);
    int j1 = 0;
    for (int dx = -2; dx <= 3; dx++) { /* vertical 6-tap of each column, then horizontal */
        int col = tap6(S(dx, -2), S(dx, -1), S(dx, 0), S(dx, 1), S(dx, 2), S(dx, 3));
        static const int w6[6] = {1, -5, 20, 20, -5, 1};
        j1 += w6[dx + 2] * col;
    }
    int j = CLIP1((j1 + 512) >> 10);
#undef S
    switch (xF * 4 + yF) { /* Table 8-12 */
    case 0 * 4 + 1: return (G + h + 1) >> 1;  /* d */
    case 0 * 4 + 2: return h;
    case 0 * 4 + 3: return (M + h + 1) >> 1;  /* n */
    case 1 * 4 + 0: return (G + b + 1) >> 1;  /* a */
    case 1 * 4 + 1: return (b + h + 1) >> 1;  /* e */
    case 1 * 4 + 2: return (h + j + 1) >> 1;  /* i */
    case 1 * 4 + 3: return (h + s + 1) >> 1;  /* p */
    case 2 * 4 + 0: return b;
    case 2 * 4 + 1: return (b + j + 1) >> 1;  /* f */
    case 2 * 4 + 2: return j;
    case 2 * 4 + 3: return (j + s + 1) >> 1;  /* q */
    case 3 * 4 + 0: return (H + b + 1) >> 1;  /* c */
    case 3 * 4 + 1: return (b + m + 1) >> 1;  /* g */
    case 3 * 4 + 2: return (j + m + 1) >> 1;  /* k */
    default: return (m + s + 1) >> 1;         /* r */
    }
}

/* 8.4.2.2.2 (H264InterPrediction.cpp:2485-2522) */
int refcpu_chroma_sample(const uint8_t *ref, int w, int h, int xI, int yI, int xF, int yF) {
    int A = px(ref, w, h, xI, yI), B = px(ref, w, h, xI + 1, yI);
    int C = px(ref, w, h, xI, yI + 1), D = px(ref, w, h, xI + 1, yI + 1);
    return ((8 - xF) * (8 - yF) * A + xF * (8 - yF) * B + (8 - xF) * yF * C + xF * yF * D + 32) >> 6;
}

/* ------------------------------------------------ picture context */
typedef struct {
    int mode;            /* 0 default, 1 weighted (explicit / implicit) */
    int logWD[3], w0[3], w1[3], o0[3], o1[3];
} Weights;

typedef struct {
    int cls;             /* H264G_MB_* */
    int qp;
    int t8;
    uint16_t nz4;        /* TotalCoeff > 0 per luma4x4BlkIdx (4x4-transform MBs) */
    uint8_t nz8;         /* non-zero levels per 8x8 block (8x8-transform MBs) */
    int refid[2][16];    /* reference picture identity per 4x4 (raster), -1 = none */
    int pred[2][16];     /* PredFlagLX per 4x4 */
    int mv[2][16][2];
    int field;           /* mb_field_decoding_flag (MBAFF frames) */
} MbInfo;

typedef struct {
    const h264g_pic_desc *d;
    const h264g_mb *mbs;
    int W, H;            /* in MBs */
    int wl, hl, wc, hc;  /* plane sizes */
    uint8_t *y, *cb, *cr;
    uint8_t *const *planes;
    MbInfo *info;
    int mbaff;           /* MbaffFrameFlag: MB addresses run over pairs, top then bottom */
} Ctx;

static int is_intra_cls(int c) { return c == H264G_MB_I4x4 || c == H264G_MB_I8x8 || c == H264G_MB_I16x16; }

/* ------------------------------------------------- MB addressing */
static int mb_field(const Ctx *X, int addr) { return X->mbaff && (X->mbs[addr].flags & H264G_MBF_FIELD) ? 1 : 0; }

/* 6.4.1 (H264PictureBase.cpp:2503-2530): luma position of the MB's first sample; a
 * field MB of an MBAFF frame starts on row 32 * pairRow + bottom. */
static void mb_luma_pos(const Ctx *X, int addr, int *x, int *y) {
    if (!X->mbaff) {
        *x = (addr % X->W) * 16;
        *y = (addr / X->W) * 16;
        return;
    }
    int pair = addr / 2;
    *x = (pair % X->W) * 16;
    *y = (pair / X->W) * 32 + (mb_field(X, addr) ? addr % 2 : (addr % 2) * 16);
}

/* Position and row pitch of the MB in a plane (n = 16 luma, 8 chroma): chroma rows of
 * MB pairs as in cSC (H264PictureBase.cpp:2133): (yL >> 4) * 8 + yL % 2. */
static void mb_plane_origin(const Ctx *X, int addr, int n, int stride, int *x0, int *y0, int *pitch) {
    int xL, yL;
    mb_luma_pos(X, addr, &xL, &yL);
    *x0 = n == 16 ? xL : (xL >> 4) * 8;
    *y0 = n == 16 ? yL : ((yL >> 4) * 8) + (yL % 2);
    *pitch = stride * (mb_field(X, addr) ? 2 : 1);
}

/* 6.4.12.2, Table 6-4 (H264PictureBase.cpp:2984-3395): the MB covering (xN, yN) relative
 * to MB addr of an MBAFF frame (-1: not available) and (xW, yW) inside it.  Pairs A, B,
 * C, D of 6.4.10 exist when inside the picture (one slice per picture). */
static int mbaff_nb(const Ctx *X, int addr, int xN, int yN, int maxW, int maxH, int *xW, int *yW) {
    int W = X->W, pair = addr / 2, top = addr % 2 == 0, curFrame = !mb_field(X, addr);
    int pA = pair % W ? 2 * (pair - 1) : -1, pB = pair >= W ? 2 * (pair - W) : -1;
    int pC = (pair >= W && (pair + 1) % W) ? 2 * (pair - W + 1) : -1, pD = (pair >= W && pair % W) ? 2 * (pair - W - 1) : -1;
    int n = -1, yM = 0;
    if (yN > maxH - 1 || (xN > maxW - 1 && yN >= 0)) return -1;
    if (xN >= 0 && yN >= 0) { n = addr; yM = yN; }
    else if (xN < 0 && yN >= 0) {
        if (pA < 0) return -1;
        int aFrame = !mb_field(X, pA);
        if (curFrame && aFrame) { n = pA + !top; yM = yN; }
        else if (curFrame) { n = pA + (yN % 2); yM = (yN + (top ? 0 : maxH)) >> 1; }
        else if (aFrame) {
            int r = (yN << 1) + !top;            /* frame row of the left pair */
            n = r < maxH ? pA : pA + 1;
            yM = r < maxH ? r : r - maxH;
        } else { n = pA + !top; yM = yN; }
    } else { /* yN < 0 */
        int pX = xN < 0 ? pD : (xN < maxW ? pB : pC);
        if (curFrame && !top) {
            if (xN > maxW - 1) return -1;
            if (xN >= 0) { n = addr - 1; yM = yN; }
            else if (pA < 0) return -1;
            else if (!mb_field(X, pA)) { n = pA; yM = yN; }
            else { n = pA + 1; yM = (yN + maxH) >> 1; }
        } else {
            if (pX < 0) return -1;
            if (curFrame || !top) { n = pX + 1; yM = yN; }
            else if (!mb_field(X, pX)) { n = pX + 1; yM = 2 * yN; }
            else { n = pX; yM = yN; }
        }
    }
    *xW = (xN + maxW) % maxW;
    *yW = (yM + maxH) % maxH;
    return n;
}

/* 6.4.12.1 (H264PictureBase.cpp:2878-2979) for non-MBAFF frames. */
static int frame_nb(const Ctx *X, int addr, int xN, int yN, int n, int *xW, int *yW) {
    int mx = addr % X->W, my = addr / X->W;
    if (xN < 0 && yN < 0) { mx--; my--; }
    else if (xN < 0 && yN < n) mx--;
    else if (xN < 0) return -1;
    else if (xN < n && yN < 0) my--;
    else if (xN < n && yN < n) { /* current MB */ }
    else if (yN < 0) { mx++; my--; }
    else return -1;
    if (mx < 0 || my < 0 || mx >= X->W || my >= X->H) return -1;
    if (X->mbs[my * X->W + mx].slice_num != X->mbs[addr].slice_num) return -1;  /* another slice */
    *xW = (xN + n) % n;
    *yW = (yN + n) % n;
    return my * X->W + mx;
}

/* ----------------------------------------------------------- levels */
typedef struct {
    int32_t dc16[16];
    int32_t l4[16][16];  /* scan order; I16x16 AC at 1..15 */
    int32_t l8[4][64];
    int32_t cdc[2][4];
    int32_t cac[2][4][16];
    int16_t pcm[384];
} Levels;

static void parse_levels(const h264g_mb *mb, const int16_t *lv, Levels *L) {
    memset(L, 0, sizeof(*L));
    const int16_t *p = lv + mb->level_offset;
    if (mb->mb_type == H264G_MB_IPCM) { memcpy(L->pcm, p, sizeof(L->pcm)); return; }
    int cbp_l = mb->cbp & 15, cbp_c = mb->cbp >> 4, i16 = mb->mb_type == H264G_MB_I16x16, t8 = mb->flags & H264G_MBF_T8x8;
    int32_t *dst[64];
    int nb = 0;
    if (i16) dst[nb++] = L->dc16;
    for (int q = 0; q < 4; q++) {
        if (!(cbp_l & (1 << q))) continue;
        if (t8) for (int w = 0; w < 4; w++) dst[nb++] = &L->l8[q][w * 16];
        else for (int b = 0; b < 4; b++) dst[nb++] = L->l4[q * 4 + b];
    }
    int dc_first = nb;
    if (cbp_c >= 1) { dst[nb++] = L->cdc[0]; dst[nb++] = L->cdc[1]; }
    int dc_end = nb;
    if (cbp_c == 2) for (int c = 0; c < 2; c++) for (int b = 0; b < 4; b++) dst[nb++] = L->cac[c][b];
    const int16_t *v = p + nb;
    for (int i = 0; i < nb; i++) {
        uint16_t m = (uint16_t)p[i];
        int n = (i >= dc_first && i < dc_end) ? 4 : 16;
        for (int k = 0; k < n; k++) if (m & (1u << k)) dst[i][k] = *v++;
    }
}

/* --------------------------------------------- weights (8.4.3 / 8.4.2.3) */
/* H264InterPrediction.cpp:2526-2613 (mode) and 2833-3050 (values). */
static int derive_weights(const h264g_pic_desc *d, int refIdxL0, int refIdxL1, int pf0, int pf1, Weights *w) {
    memset(w, 0, sizeof(*w));
    int st = d->slice_type;
    if (st == H264G_SLICE_P) w->mode = d->weighted_pred_flag == 1;
    else if (st == H264G_SLICE_B) w->mode = d->weighted_bipred_idc == 1 || (d->weighted_bipred_idc == 2 && pf0 && pf1);
    if (!w->mode) return 0;
    if (st == H264G_SLICE_B && d->weighted_bipred_idc == 2) {
        /* implicit: POC distances; tb/td clipped (8-201, 8-202) */
        int poc0 = d->ref_poc[0][refIdxL0], poc1 = d->ref_poc[1][refIdxL1];
        int diff10 = poc1 - poc0;
        int ww0 = 32, ww1 = 32;
        if (diff10 != 0 && !d->ref_long_term[0][refIdxL0] && !d->ref_long_term[1][refIdxL1]) {
            /* With POC1 == POC0 the reference divides by td = 0 (undefined); the spec's
             * 32/32 fallback applies there and is what this restatement returns. */
            int tb = CLIP3(-128, 127, d->poc - poc0), td = CLIP3(-128, 127, diff10);
            int tx = (16384 + iabs(td / 2)) / td;
            int dsf = CLIP3(-1024, 1023, (tb * tx + 32) >> 6);
            if (!((dsf >> 2) < -64 || (dsf >> 2) > 128)) { ww0 = 64 - (dsf >> 2); ww1 = dsf >> 2; }
        }
        for (int c = 0; c < 3; c++) { w->logWD[c] = 5; w->w0[c] = ww0; w->w1[c] = ww1; }
        return 0;
    }
    /* explicit: index -1 of the L0/L1 tables aliases the previous member (Q2) */
    int i0 = refIdxL0, i1 = refIdxL1;
    w->logWD[0] = d->luma_log2_weight_denom;
    w->logWD[1] = w->logWD[2] = d->chroma_log2_weight_denom;
    w->w0[0] = i0 >= 0 ? d->luma_weight[0][i0] : d->luma_weight_flag_last[0];
    w->o0[0] = i0 >= 0 ? d->luma_offset[0][i0] : d->luma_weight[0][31];
    w->w1[0] = i1 >= 0 ? d->luma_weight[1][i1] : d->luma_weight_flag_last[1];
    w->o1[0] = i1 >= 0 ? d->luma_offset[1][i1] : d->luma_weight[1][31];
    for (int c = 0; c < 2; c++) {
        w->w0[1 + c] = i0 >= 0 ? d->chroma_weight[0][i0][c] : 0;
        w->o0[1 + c] = i0 >= 0 ? d->chroma_offset[0][i0][c] : 0;
        w->w1[1 + c] = i1 >= 0 ? d->chroma_weight[1][i1][c] : 0;
        w->o1[1 + c] = i1 >= 0 ? d->chroma_offset[1][i1][c] : 0;
    }
    if (pf0 && pf1) { /* bipred weight-sum constraint (H264InterPrediction.cpp:3038-3047) */
        if (!(-128 <= w->w0[0] + w->w1[0] && w->w0[0] + w->w1[0] <= (w->logWD[0] == 7 ? 127 : 128))) return H264G_E_SYNTAX;
        if (!(-128 <= w->w0[1] + w->w1[1] && w->w0[1] + w->w1[1] <= (w->logWD[1] == 7 ? 127 : 128))) return H264G_E_SYNTAX;
        if (!(-128 <= w->w0[2] + w->w1[2] && w->w0[1] + w->w1[2] <= (w->logWD[2] == 7 ? 127 : 128))) return H264G_E_SYNTAX;
    }
    return 0;
}

/* 8.4.2.3.1 / 8.4.2.3.2 for component c (0 = Y); L1-only luma uses w0/o0 (Q2). */
static int weigh(const Weights *w, int c, int pf0, int pf1, int p0, int p1) {
    if (!w->mode) {
        if (pf0 && pf1) return (p0 + p1 + 1) >> 1;
        return pf0 ? p0 : p1;
    }
    int lw = w->logWD[c];
    if (pf0 && pf1) return CLIP1(((p0 * w->w0[c] + p1 * w->w1[c] + (1 << lw)) >> (lw + 1)) + ((w->o0[c] + w->o1[c] + 1) >> 1));
    int p = pf0 ? p0 : p1;
    int ww, oo;
    if (pf0) { ww = w->w0[c]; oo = w->o0[c]; }
    else if (c == 0) { ww = w->w0[0]; oo = w->o0[0]; }  /* H264InterPrediction.cpp:2764,2768 */
    else { ww = w->w1[c]; oo = w->o1[c]; }
    if (lw >= 1) return CLIP1(((p * ww + (1 << (lw - 1))) >> lw) + oo);
    return CLIP1(p * ww + oo);
}

/* -------------------------------------------------- inter MB (8.4) */
/* Per 4x4 block: reference selection (8.4.2.1, H264InterPrediction.cpp:2117-2224),
 * chroma vector (8.4.1.4, :2006-2047), interpolation (8.4.2.2) and weighting (8.4.2.3).
 * A field MB of an MBAFF frame predicts from the field of RefPicListX[refIdx / 2] with
 * the MB's parity when refIdx is even, the other one when odd; that field is the frame
 * buffer seen from row `parity` with twice the row pitch, and the reference's field view
 * reports that pitch as its width, which is also where it clamps x (Q17).  Its weights
 * are looked up at refIdx >> 1 (H264InterPrediction.cpp:2993-3031). */
static int inter_mb(Ctx *X, int addr, const h264g_mb *mb, MbInfo *info) {
    const h264g_pic_desc *d = X->d;
    const int field = mb_field(X, addr), bottom = X->mbaff && addr % 2;
    int x0, y0, pitch, cx0, cy0, cpitch;
    mb_plane_origin(X, addr, 16, X->wl, &x0, &y0, &pitch);
    mb_plane_origin(X, addr, 8, X->wc, &cx0, &cy0, &cpitch);
    uint8_t *oy = X->y + y0 * X->wl + x0, *ocb = X->cb + cy0 * X->wc + cx0, *ocr = X->cr + cy0 * X->wc + cx0;
    /* MB origin in the coordinates of the reference picture it predicts from: a field MB
     * sits at field row 16 * pairRow (H264InterPrediction.cpp:553-557) */
    const int xMb = x0, yMb = field ? (y0 - bottom) / 2 : y0;
    for (int b = 0; b < 16; b++) {
        int bx = b % 4, by = b / 4, q = (by / 2) * 2 + bx / 2;
        int ref[2], pf[2];
        for (int l = 0; l < 2; l++) {
            ref[l] = mb->ref_idx[l][q];
            pf[l] = ref[l] >= 0;
            int frame_idx = field ? ref[l] >> 1 : ref[l];
            if (pf[l] && (frame_idx >= d->num_ref_idx_active[l] || d->ref_slot[l][frame_idx] < 0)) return H264G_E_ARG;
        }
        if (!pf[0] && !pf[1]) return H264G_E_ARG;
        if (field && d->slice_type == H264G_SLICE_B && d->weighted_bipred_idc == 2 && pf[0] && pf[1])
            return H264G_E_ARG; /* implicit weights of field MBs need field POCs */
        Weights w;
        int rc = derive_weights(d, field ? ref[0] >> 1 : ref[0], field ? ref[1] >> 1 : ref[1], pf[0], pf[1], &w);
        if (rc) return rc;
        int pl[2][16], pcb[2][4], pcr[2][4];
        for (int l = 0; l < 2; l++) {
            if (!pf[l]) continue;
            const uint8_t *rp = X->planes[d->ref_slot[l][field ? ref[l] >> 1 : ref[l]]];
            const uint8_t *ry = rp, *rcb = rp + X->wl * X->hl, *rcr = rcb + X->wc * X->hc;
            int wl = X->wl, hl = X->hl, wc = X->wc, hc = X->hc, mvcy_off = 0;
            int mvx = mb->mv[l][b][0], mvy = mb->mv[l][b][1];
            if (field) {
                int parity = (ref[l] & 1) ^ bottom;  /* 0: top field, 1: bottom field */
                ry += parity * wl; rcb += parity * wc; rcr += parity * wc;
                wl *= 2; wc *= 2; hl /= 2; hc /= 2;
                if (parity && !bottom) mvcy_off = -2;      /* Table 8-10 */
                else if (!parity && bottom) mvcy_off = 2;
            }
            int xAL = xMb + bx * 4, yAL = yMb + by * 4;
            for (int yy = 0; yy < 4; yy++)
                for (int xx = 0; xx < 4; xx++)
                    pl[l][yy * 4 + xx] = refcpu_luma_sample(ry, wl, hl, xAL + (mvx >> 2) + xx, yAL + (mvy >> 2) + yy, mvx & 3, mvy & 3);
            int mvcy = mvy + mvcy_off;
            for (int yy = 0; yy < 2; yy++)
                for (int xx = 0; xx < 2; xx++) {
                    int xi = xAL / 2 + (mvx >> 3) + xx, yi = yAL / 2 + (mvcy >> 3) + yy;
                    pcb[l][yy * 2 + xx] = refcpu_chroma_sample(rcb, wc, hc, xi, yi, mvx & 7, mvcy & 7);
                    pcr[l][yy * 2 + xx] = refcpu_chroma_sample(rcr, wc, hc, xi, yi, mvx & 7, mvcy & 7);
                }
        }
        for (int yy = 0; yy < 4; yy++)
            for (int xx = 0; xx < 4; xx++)
                oy[(by * 4 + yy) * pitch + bx * 4 + xx] = (uint8_t)weigh(&w, 0, pf[0], pf[1], pl[0][yy * 4 + xx], pl[1][yy * 4 + xx]);
        for (int yy = 0; yy < 2; yy++)
            for (int xx = 0; xx < 2; xx++) {
                int o = (by * 2 + yy) * cpitch + bx * 2 + xx;
                ocb[o] = (uint8_t)weigh(&w, 1, pf[0], pf[1], pcb[0][yy * 2 + xx], pcb[1][yy * 2 + xx]);
                ocr[o] = (uint8_t)weigh(&w, 2, pf[0], pf[1], pcr[0][yy * 2 + xx], pcr[1][yy * 2 + xx]);
            }
        for (int l = 0; l < 2; l++) {
            info->pred[l][b] = pf[l];
            /* bS identity: RefPicListX[RefIdxLX], the field index for field MBs (Q4) */
            info->refid[l][b] = pf[l] ? d->ref_slot[l][ref[l]] : -1;
            info->mv[l][b][0] = mb->mv[l][b][0];
            info->mv[l][b][1] = mb->mv[l][b][1];
        }
    }
    return 0;
}

/* luma4x4BlkIdx -> position inside the MB (6.4.3) */
static void blk_xy(int b, int *x, int *y) {
    *x = ((b >> 2) & 1) * 8 + (b & 1) * 4;
    *y = ((b >> 3) & 1) * 8 + ((b >> 1) & 1) * 4;
}

static void add_residual(uint8_t *o, int pitch, int x, int y, const int32_t *r, int n) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            uint8_t *s = &o[(y + i) * pitch + x + j];
            *s = (uint8_t)CLIP1(*s + r[i * n + j]);
        }
}

/* residual of the luma 4x4 block b for non-I16x16 MBs (8.5.1); field MBs use the field
 * scan (field_pic_flag | mb_field_decoding_flag, H264PictureBase.cpp:3419) */
static void luma4x4_residual(const Levels *L, int b, int qp, int field, const uint8_t *w, int32_t r[16]) {
    residual4x4_w(L->l4[b], qp, 0, field, w, r);
}

/* chroma residual of one component: rMb[8][8] raster (8.5.4); w: the MB's luma 4x4 list
 * (quirk Q13: the chroma drivers never derive their own LevelScale4x4,
 * H264PictureBase.cpp:3752-3929) */
static void chroma_residual(const Levels *L, int c, int qpc, int field, const uint8_t *w, int32_t rmb[64]) {
    int32_t dcC[4];
    chroma_dc_w(L->cdc[c], qpc, w, dcC);
    for (int b = 0; b < 4; b++) {
        int32_t list[16], r[16];
        memcpy(list, L->cac[c][b], sizeof(list));
        list[0] = dcC[b];
        residual4x4_w(list, qpc, 1, field, w, r);
        int x0 = (b % 2) * 4, y0 = (b / 2) * 4;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) rmb[(y0 + i) * 8 + x0 + j] = r[i * 4 + j];
    }
}

/* ------------------------------------------------ intra (8.3) */
/* Neighbouring sample (xN, yN) relative to the MB's top-left in a plane with MB size
 * n (16 luma, 8 chroma): the MB covering it by 6.4.12 (Table 6-3 / 6-4), -1 if not
 * available (outside the picture, or an inter MB under constrained intra), else the
 * sample at that MB's position (6.4.1) plus (xW, yW), rows two apart in a field MB of
 * an MBAFF frame (H264PictureBase.cpp:1099-1155, 1917-1925, 2172-2186). */
static int nb_sample(Ctx *X, const uint8_t *plane, int stride, int n, int addr, int xN, int yN) {
    int xW = 0, yW = 0;
    int na = X->mbaff ? mbaff_nb(X, addr, xN, yN, n, n, &xW, &yW) : frame_nb(X, addr, xN, yN, n, &xW, &yW);
    if (na < 0) return -1;
    if (X->d->constrained_intra_pred_flag) {  /* quirk Q18 */
        const h264g_mb *m = &X->mbs[na];
        if (m->mb_type == H264G_MB_P_SKIP ||
            ((m->mb_type == H264G_MB_P || m->mb_type == H264G_MB_B) && m->part_shape != H264G_PART_8x8))
            return -1;
    }
    int x0, y0, pitch;
    mb_plane_origin(X, na, n, stride, &x0, &y0, &pitch);
    return plane[y0 * stride + yW * pitch + x0 + xW];
}

#define P4(x, y) p[((y) + 1) * 9 + (x) + 1]
/* 8.3.1.2 (H264PictureBase.cpp:1062-1419) for the 4x4 block at (xO, yO). */
static int intra4x4_pred(Ctx *X, int addr, int blk, int mode, uint8_t out[16]) {
    int xO, yO, p[45];
    blk_xy(blk, &xO, &yO);
    for (int i = 0; i < 45; i++) p[i] = -1;
    P4(-1, -1) = nb_sample(X, X->y, X->wl, 16, addr, xO - 1, yO - 1);
    for (int y = 0; y < 4; y++) P4(-1, y) = nb_sample(X, X->y, X->wl, 16, addr, xO - 1, yO + y);
    for (int x = 0; x < 8; x++)
        P4(x, -1) = (x > 3 && (blk == 3 || blk == 11)) ? -1 : nb_sample(X, X->y, X->wl, 16, addr, xO + x, yO - 1);
    if (P4(4, -1) < 0 && P4(5, -1) < 0 && P4(6, -1) < 0 && P4(7, -1) < 0 && P4(3, -1) >= 0)
        for (int x = 4; x < 8; x++) P4(x, -1) = P4(3, -1);
    int top = P4(0, -1) >= 0 && P4(1, -1) >= 0 && P4(2, -1) >= 0 && P4(3, -1) >= 0;
    int left = P4(-1, 0) >= 0 && P4(-1, 1) >= 0 && P4(-1, 2) >= 0 && P4(-1, 3) >= 0;
    int tl = P4(-1, -1) >= 0;
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) {
            int v;
            switch (mode) {
            case 0: if (!top) return H264G_E_SYNTAX; v = P4(x, -1); break;
            case 1: if (!left) return H264G_E_SYNTAX; v = P4(-1, y); break;
            case 2:
                if (top && left) v = (P4(0, -1) + P4(1, -1) + P4(2, -1) + P4(3, -1) + P4(-1, 0) + P4(-1, 1) + P4(-1, 2) + P4(-1, 3) + 4) >> 3;
                else if (left) v = (P4(-1, 0) + P4(-1, 1) + P4(-1, 2) + P4(-1, 3) + 2) >> 2;
                else if (top) v = (P4(0, -1) + P4(1, -1) + P4(2, -1) + P4(3, -1) + 2) >> 2;
                else v = 128;
                break;
            case 3:
                if (!top) return H264G_E_SYNTAX;
                v = (x == 3 && y == 3) ? (P4(6, -1) + 3 * P4(7, -1) + 2) >> 2
                                       : (P4(x + y, -1) + 2 * P4(x + y + 1, -1) + P4(x + y + 2, -1) + 2) >> 2;
                break;
            case 4:
                if (!(top && left && tl)) return H264G_E_SYNTAX;
                if (x > y) v = (P4(x - y - 2, -1) + 2 * P4(x - y - 1, -1) + P4(x - y, -1) + 2) >> 2;
                else if (x < y) v = (P4(-1, y - x - 2) + 2 * P4(-1, y - x - 1) + P4(-1, y - x) + 2) >> 2;
                else v = (P4(0, -1) + 2 * P4(-1, -1) + P4(-1, 0) + 2) >> 2;
                break;
            case 5: {
                if (!(top && left && tl)) return H264G_E_SYNTAX;
                int z = 2 * x - y;
                if (z >= 0 && (z & 1) == 0) v = (P4(x - (y >> 1) - 1, -1) + P4(x - (y >> 1), -1) + 1) >> 1;
                else if (z >= 0) v = (P4(x - (y >> 1) - 2, -1) + 2 * P4(x - (y >> 1) - 1, -1) + P4(x - (y >> 1), -1) + 2) >> 2;
                else if (z == -1) v = (P4(-1, 0) + 2 * P4(-1, -1) + P4(0, -1) + 2) >> 2;
                else v = (P4(-1, y - 1) + 2 * P4(-1, y - 2) + P4(-1, y - 3) + 2) >> 2;
                break;
            }
            case 6: {
                if (!(top && left && tl)) return H264G_E_SYNTAX;
                int z = 2 * y - x;
                if (z >= 0 && (z & 1) == 0) v = (P4(-1, y - (x >> 1) - 1) + P4(-1, y - (x >> 1)) + 1) >> 1;
                else if (z >= 0) v = (P4(-1, y - (x >> 1) - 2) + 2 * P4(-1, y - (x >> 1) - 1) + P4(-1, y - (x >> 1)) + 2) >> 2;
                else if (z == -1) v = (P4(-1, 0) + 2 * P4(-1, -1) + P4(0, -1) + 2) >> 2;
                else v = (P4(x - 1, -1) + 2 * P4(x - 2, -1) + P4(x - 3, -1) + 2) >> 2;
                break;
            }
            case 7:
                if (!top) return H264G_E_SYNTAX;
                if ((y & 1) == 0) v = (P4(x + (y >> 1), -1) + P4(x + (y >> 1) + 1, -1) + 1) >> 1;
                else v = (P4(x + (y >> 1), -1) + 2 * P4(x + (y >> 1) + 1, -1) + P4(x + (y >> 1) + 2, -1) + 2) >> 2;
                break;
            default: {
                if (!left) return H264G_E_SYNTAX;
                int z = x + 2 * y;
                if (z < 5 && (z & 1) == 0) v = (P4(-1, y + (x >> 1)) + P4(-1, y + (x >> 1) + 1) + 1) >> 1;
                else if (z < 5) v = (P4(-1, y + (x >> 1)) + 2 * P4(-1, y + (x >> 1) + 1) + P4(-1, y + (x >> 1) + 2) + 2) >> 2;
                else if (z == 5) v = (P4(-1, 2) + 3 * P4(-1, 3) + 2) >> 2;
                else v = P4(-1, 3);
                break;
            }
            }
            out[y * 4 + x] = (uint8_t)v;
        }
    return 0;
}
#undef P4

#define P8(x, y) p[((y) + 1) * 17 + (x) + 1]
#define Q8(x, y) q[((y) + 1) * 17 + (x) + 1]
/* 8.3.2.2 incl. reference sample filtering 8.3.2.2.1 (H264PictureBase.cpp:1423-1843) */
static int intra8x8_pred(Ctx *X, int addr, int blk, int mode, uint8_t out[64]) {
    int xO = (blk % 2) * 8, yO = (blk / 2) * 8, p[9 * 17], q[9 * 17];
    for (int i = 0; i < 9 * 17; i++) p[i] = q[i] = -1;
    P8(-1, -1) = nb_sample(X, X->y, X->wl, 16, addr, xO - 1, yO - 1);
    for (int y = 0; y < 8; y++) P8(-1, y) = nb_sample(X, X->y, X->wl, 16, addr, xO - 1, yO + y);
    for (int x = 0; x < 16; x++) P8(x, -1) = nb_sample(X, X->y, X->wl, 16, addr, xO + x, yO - 1);
    int up8 = 1;
    for (int x = 8; x < 16; x++) up8 &= P8(x, -1) < 0;
    if (up8 && P8(7, -1) >= 0) for (int x = 8; x < 16; x++) P8(x, -1) = P8(7, -1);
    int top16 = 1, left8 = 1;
    for (int x = 0; x < 16; x++) top16 &= P8(x, -1) >= 0;
    for (int y = 0; y < 8; y++) left8 &= P8(-1, y) >= 0;
    if (top16) {
        Q8(0, -1) = P8(-1, -1) >= 0 ? (P8(-1, -1) + 2 * P8(0, -1) + P8(1, -1) + 2) >> 2 : (3 * P8(0, -1) + P8(1, -1) + 2) >> 2;
        for (int x = 1; x < 15; x++) Q8(x, -1) = (P8(x - 1, -1) + 2 * P8(x, -1) + P8(x + 1, -1) + 2) >> 2;
        Q8(15, -1) = (P8(14, -1) + 3 * P8(15, -1) + 2) >> 2;
    }
    if (P8(-1, -1) >= 0) {
        if (P8(0, -1) < 0 || P8(-1, 0) < 0) {
            if (P8(0, -1) >= 0) Q8(-1, -1) = (3 * P8(-1, -1) + P8(0, -1) + 2) >> 2;
            else if (P8(-1, 0) >= 0) Q8(-1, -1) = (3 * P8(-1, -1) + P8(-1, 0) + 2) >> 2;
            else Q8(-1, -1) = P8(-1, -1);
        } else Q8(-1, -1) = (P8(0, -1) + 2 * P8(-1, -1) + P8(-1, 0) + 2) >> 2;
    }
    if (left8) {
        Q8(-1, 0) = P8(-1, -1) >= 0 ? (P8(-1, -1) + 2 * P8(-1, 0) + P8(-1, 1) + 2) >> 2 : (3 * P8(-1, 0) + P8(-1, 1) + 2) >> 2;
        for (int y = 1; y < 7; y++) Q8(-1, y) = (P8(-1, y - 1) + 2 * P8(-1, y) + P8(-1, y + 1) + 2) >> 2;
        Q8(-1, 7) = (P8(-1, 6) + 3 * P8(-1, 7) + 2) >> 2;
    }
    /* predictions use the filtered samples only */
    int top = 1, left = left8, tl = Q8(-1, -1) >= 0;
    for (int x = 0; x < 8; x++) top &= Q8(x, -1) >= 0;
    int top16f = top16;
#define F(x, y) Q8(x, y)
    for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) {
            int v;
            switch (mode) {
            case 0: if (!top) return H264G_E_SYNTAX; v = F(x, -1); break;
            case 1: if (!left) return H264G_E_SYNTAX; v = F(-1, y); break;
            case 2: {
                int st = 0, sl = 0;
                for (int k = 0; k < 8; k++) { st += F(k, -1); sl += F(-1, k); }
                if (top && left) v = (st + sl + 8) >> 4;
                else if (left) v = (sl + 4) >> 3;
                else if (top) v = (st + 4) >> 3;
                else v = 128;
                break;
            }
            case 3:
                if (!top16f) return H264G_E_SYNTAX;
                v = (x == 7 && y == 7) ? (F(14, -1) + 3 * F(15, -1) + 2) >> 2
                                       : (F(x + y, -1) + 2 * F(x + y + 1, -1) + F(x + y + 2, -1) + 2) >> 2;
                break;
            case 4:
                if (!(top && left && tl)) return H264G_E_SYNTAX;
                if (x > y) v = (F(x - y - 2, -1) + 2 * F(x - y - 1, -1) + F(x - y, -1) + 2) >> 2;
                else if (x < y) v = (F(-1, y - x - 2) + 2 * F(-1, y - x - 1) + F(-1, y - x) + 2) >> 2;
                else v = (F(0, -1) + 2 * F(-1, -1) + F(-1, 0) + 2) >> 2;
                break;
            case 5: {
                if (!(top && left && tl)) return H264G_E_SYNTAX;
                int z = 2 * x - y;
                if (z >= 0 && (z & 1) == 0) v = (F(x - (y >> 1) - 1, -1) + F(x - (y >> 1), -1) + 1) >> 1;
                else if (z >= 0) v = (F(x - (y >> 1) - 2, -1) + 2 * F(x - (y >> 1) - 1, -1) + F(x - (y >> 1), -1) + 2) >> 2;
                else if (z == -1) v = (F(-1, 0) + 2 * F(-1, -1) + F(0, -1) + 2) >> 2;
                else v = (F(-1, y - 2 * x - 1) + 2 * F(-1, y - 2 * x - 2) + F(-1, y - 2 * x - 3) + 2) >> 2;
                break;
            }
            case 6: {
                if (!(top && left && tl)) return H264G_E_SYNTAX;
                int z = 2 * y - x;
                if (z >= 0 && (z & 1) == 0) v = (F(-1, y - (x >> 1) - 1) + F(-1, y - (x >> 1)) + 1) >> 1;
                else if (z >= 0) v = (F(-1, y - (x >> 1) - 2) + 2 * F(-1, y - (x >> 1) - 1) + F(-1, y - (x >> 1)) + 2) >> 2;
                else if (z == -1) v = (F(-1, 0) + 2 * F(-1, -1) + F(0, -1) + 2) >> 2;
                else v = (F(x - 2 * y - 1, -1) + 2 * F(x - 2 * y - 2, -1) + F(x - 2 * y - 3, -1) + 2) >> 2;
                break;
            }
            case 7:
                if (!top16f) return H264G_E_SYNTAX;
                if ((y & 1) == 0) v = (F(x + (y >> 1), -1) + F(x + (y >> 1) + 1, -1) + 1) >> 1;
                else v = (F(x + (y >> 1), -1) + 2 * F(x + (y >> 1) + 1, -1) + F(x + (y >> 1) + 2, -1) + 2) >> 2;
                break;
            default: {
                if (!left) return H264G_E_SYNTAX;
                int z = x + 2 * y;
                if (z < 13 && (z & 1) == 0) v = (F(-1, y + (x >> 1)) + F(-1, y + (x >> 1) + 1) + 1) >> 1;
                else if (z < 13) v = (F(-1, y + (x >> 1)) + 2 * F(-1, y + (x >> 1) + 1) + F(-1, y + (x >> 1) + 2) + 2) >> 2;
                else if (z == 13) v = (F(-1, 6) + 3 * F(-1, 7) + 2) >> 2;
                else v = F(-1, 7);
                break;
            }
            }
            out[y * 8 + x] = (uint8_t)v;
        }
#undef F
    return 0;
}
#undef P8
#undef Q8

/* 8.3.3 (H264PictureBase.cpp:1847-2056) */
static int intra16_pred(Ctx *X, int addr, int mode, uint8_t out[256]) {
    int top[16], left[16], tl, ta = 1, la = 1;
    for (int i = 0; i < 16; i++) {
        top[i] = nb_sample(X, X->y, X->wl, 16, addr, i, -1);
        left[i] = nb_sample(X, X->y, X->wl, 16, addr, -1, i);
        ta &= top[i] >= 0; la &= left[i] >= 0;
    }
    tl = nb_sample(X, X->y, X->wl, 16, addr, -1, -1);
    int dc = 128;
    if (mode == 2) {
        int st = 0, sl = 0;
        for (int i = 0; i < 16; i++) { st += top[i]; sl += left[i]; }
        if (ta && la) dc = (st + sl + 16) >> 5;
        else if (la) dc = (sl + 8) >> 4;
        else if (ta) dc = (st + 8) >> 4;
    }
    int a = 0, b = 0, c = 0;
    if (mode == 3) {
        if (!(ta && la)) return H264G_E_SYNTAX;
        int H = 0, V = 0;
        for (int x = 0; x < 8; x++) H += (x + 1) * (top[8 + x] - (6 - x >= 0 ? top[6 - x] : tl));
        for (int y = 0; y < 8; y++) V += (y + 1) * (left[8 + y] - (6 - y >= 0 ? left[6 - y] : tl));
        a = 16 * (left[15] + top[15]);
        b = (5 * H + 32) >> 6;
        c = (5 * V + 32) >> 6;
    }
    if ((mode == 0 && !ta) || (mode == 1 && !la)) return H264G_E_SYNTAX;
    for (int y = 0; y < 16; y++)
        for (int x = 0; x < 16; x++) {
            int v = mode == 0 ? top[x] : mode == 1 ? left[y] : mode == 2 ? dc : CLIP1((a + b * (x - 7) + c * (y - 7) + 16) >> 5);
            out[y * 16 + x] = (uint8_t)v;
        }
    return 0;
}

/* 8.3.4 for 4:2:0 (H264PictureBase.cpp:2076-2381); DC availability uses > 0 (Q1). */
static int intra_chroma_pred(Ctx *X, const uint8_t *plane, int addr, int mode, uint8_t out[64]) {
    int top[8], left[8], tl;
    for (int i = 0; i < 8; i++) {
        top[i] = nb_sample(X, plane, X->wc, 8, addr, i, -1);
        left[i] = nb_sample(X, plane, X->wc, 8, addr, -1, i);
    }
    tl = nb_sample(X, plane, X->wc, 8, addr, -1, -1);
    if (mode == 0) {
        for (int blk = 0; blk < 4; blk++) {
            int xO = (blk % 2) * 4, yO = (blk / 2) * 4;
            int t = top[xO] > 0 && top[xO + 1] > 0 && top[xO + 2] > 0 && top[xO + 3] > 0;
            int l = left[yO] > 0 && left[yO + 1] > 0 && left[yO + 2] > 0 && left[yO + 3] > 0;
            int st = top[xO] + top[xO + 1] + top[xO + 2] + top[xO + 3];
            int sl = left[yO] + left[yO + 1] + left[yO + 2] + left[yO + 3];
            int v;
            if ((xO == 0 && yO == 0) || (xO > 0 && yO > 0)) {
                if (t && l) v = (st + sl + 4) >> 3;
                else if (l) v = (sl + 2) >> 2;
                else if (t) v = (st + 2) >> 2;
                else v = 128;
            } else if (xO > 0) { /* yO == 0 */
                v = t ? (st + 2) >> 2 : (l ? (sl + 2) >> 2 : 128);
            } else {             /* xO == 0, yO > 0 */
                v = l ? (sl + 2) >> 2 : (t ? (st + 2) >> 2 : 128);
            }
            for (int y = 0; y < 4; y++)
                for (int x = 0; x < 4; x++) out[(yO + y) * 8 + xO + x] = (uint8_t)v;
        }
        return 0;
    }
    int ta = 1, la = 1;
    for (int i = 0; i < 8; i++) { ta &= top[i] >= 0; la &= left[i] >= 0; }
    if (mode == 1) { if (!la) return H264G_E_SYNTAX; for (int y = 0; y < 8; y++) for (int x = 0; x < 8; x++) out[y * 8 + x] = (uint8_t)left[y]; return 0; }
    if (mode == 2) { if (!ta) return H264G_E_SYNTAX; for (int y = 0; y < 8; y++) for (int x = 0; x < 8; x++) out[y * 8 + x] = (uint8_t)top[x]; return 0; }
    if (!(ta && la && tl >= 0)) return H264G_E_SYNTAX;
    int H = 0, V = 0;
    for (int x = 0; x < 4; x++) H += (x + 1) * (top[4 + x] - (2 - x >= 0 ? top[2 - x] : tl));
    for (int y = 0; y < 4; y++) V += (y + 1) * (left[4 + y] - (2 - y >= 0 ? left[2 - y] : tl));
    int a = 16 * (left[7] + top[7]), b = (34 * H + 32) >> 6, c = (34 * V + 32) >> 6;
    for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) out[y * 8 + x] = (uint8_t)CLIP1((a + b * (x - 3) + c * (y - 3) + 16) >> 5);
    return 0;
}

/* Writes an n x n block at (x, y) of the MB whose first sample is o, rows `pitch` apart
 * (8.5.14, H264PictureBase.cpp:4408-4538: field MBs of MBAFF frames every other row). */
static void put_block(uint8_t *o, int pitch, int x, int y, const uint8_t *v, int n) {
    for (int i = 0; i < n; i++) memcpy(&o[(y + i) * pitch + x], &v[i * n], (size_t)n);
}

/* Intra MB reconstruction in the reference's order (H264SliceData.cpp:400-453). */
static int intra_mb(Ctx *X, int addr, const h264g_mb *mb, const Levels *L) {
    int qp = mb->qp_y, rc, field = mb_field(X, addr);
    /* scaling matrices: the intra luma lists, chroma included (Q13) */
    const uint8_t *w4 = X->d->scaling_matrix ? X->d->weight_scale4x4[0] : NULL;
    const uint8_t *w8 = X->d->scaling_matrix ? X->d->weight_scale8x8[0] : NULL;
    int x0, y0, pitch;
    mb_plane_origin(X, addr, 16, X->wl, &x0, &y0, &pitch);
    uint8_t *o = X->y + y0 * X->wl + x0;
    if (mb->mb_type == H264G_MB_I4x4) {
        for (int b = 0; b < 16; b++) {
            int32_t r[16];
            uint8_t pred[16];
            int bx, by;
            blk_xy(b, &bx, &by);
            luma4x4_residual(L, b, qp, field, w4, r);
            if ((rc = intra4x4_pred(X, addr, b, mb->intra_pred_modes[b], pred))) return rc;
            put_block(o, pitch, bx, by, pred, 4);
            add_residual(o, pitch, bx, by, r, 4);
        }
    } else if (mb->mb_type == H264G_MB_I8x8) {
        for (int b = 0; b < 4; b++) {
            int32_t r[64];
            uint8_t pred[64];
            residual8x8_w(L->l8[b], qp, field, w8, r);
            if ((rc = intra8x8_pred(X, addr, b, mb->intra_pred_modes[b], pred))) return rc;
            put_block(o, pitch, (b % 2) * 8, (b / 2) * 8, pred, 8);
            add_residual(o, pitch, (b % 2) * 8, (b / 2) * 8, r, 8);
        }
    } else {
        int32_t dcY[16], rmb[256];
        uint8_t pred[256];
        luma_dc_scan(L->dc16, qp, field, w4, dcY);
        for (int b = 0; b < 16; b++) {
            int32_t list[16], r[16];
            int bx, by;
            blk_xy(b, &bx, &by);
            memcpy(list, L->l4[b], sizeof(list));
            list[0] = dcY[(by / 4) * 4 + bx / 4];  /* dcY raster index of block b */
            residual4x4_w(list, qp, 1, field, w4, r);
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) rmb[(by + i) * 16 + bx + j] = r[i * 4 + j];
        }
        if ((rc = intra16_pred(X, addr, mb->intra16x16_pred_mode, pred))) return rc;
        put_block(o, pitch, 0, 0, pred, 16);
        add_residual(o, pitch, 0, 0, rmb, 16);
    }
    mb_plane_origin(X, addr, 8, X->wc, &x0, &y0, &pitch);
    for (int c = 0; c < 2; c++) {
        uint8_t *plane = c == 0 ? X->cb : X->cr, *oc = plane + y0 * X->wc + x0;
        int off = c == 0 ? X->d->chroma_qp_index_offset : X->d->second_chroma_qp_index_offset;
        int32_t rmb[64];
        uint8_t pred[64];
        chroma_residual(L, c, refcpu_chroma_qp(qp, off), field, w4, rmb);
        if ((rc = intra_chroma_pred(X, plane, addr, mb->intra_chroma_pred_mode, pred))) return rc;
        put_block(oc, pitch, 0, 0, pred, 8);
        add_residual(oc, pitch, 0, 0, rmb, 8);
    }
    return 0;
}

/* ---------------------------------------------- 8.7 deblocking */
/* |dx| >= 4 or |dy| >= ylim quarter samples; ylim is 2 where the q MB is a field MB of an
 * MBAFF frame (Deblock.cpp:1157-1158) */
static int mv_far(const int *a, const int *b, int ylim) { return iabs(a[0] - b[0]) >= 4 || iabs(a[1] - b[1]) >= ylim; }

/* The bS 2 test of 8.7.2.1 (Deblock.cpp:1136-1148): non-zero levels in the transform
 * block of p0 or q0. */
static int bs_nonzero(const MbInfo *P, const MbInfo *Q, int xp, int yp, int xq, int yq) {
    int bp4 = 8 * (yp / 8) + 4 * (xp / 8) + 2 * ((yp % 8) / 4) + (xp % 8) / 4;
    int bq4 = 8 * (yq / 8) + 4 * (xq / 8) + 2 * ((yq % 8) / 4) + (xq % 8) / 4;
    int bp8 = 2 * (yp / 8) + xp / 8, bq8 = 2 * (yq / 8) + xq / 8;
    return (P->t8 && (P->nz8 >> bp8 & 1)) || (!P->t8 && (P->nz4 >> bp4 & 1)) ||
           (Q->t8 && (Q->nz8 >> bq8 & 1)) || (!Q->t8 && (Q->nz4 >> bq4 & 1));
}

/* The bS 1 motion tests of 8.7.2.1 (Deblock.cpp:1170-1305) with quirk Q3. */
static int bs_motion(const MbInfo *P, const MbInfo *Q, int xp, int yp, int xq, int yq, int ylim) {
    int rp = (yp / 4) * 4 + xp / 4, rq = (yq / 4) * 4 + xq / 4;  /* raster 4x4 */
    int p0 = P->refid[0][rp], p1 = P->refid[1][rp], q0 = Q->refid[0][rq], q1 = Q->refid[1][rq];
    int np = P->pred[0][rp] + P->pred[1][rp], nq = Q->pred[0][rq] + Q->pred[1][rq];
    if (!(((p0 == q0 && p1 == q1) || (p0 == q1 && p1 == q0)) && np == nq)) return 1;
    const int *mp0 = P->mv[0][rp], *mp1 = P->mv[1][rp], *mq0 = Q->mv[0][rq], *mq1 = Q->mv[1][rq];
    int fp0 = P->pred[0][rp], fp1 = P->pred[1][rp], fq0 = Q->pred[0][rq], fq1 = Q->pred[1][rq];
    if (fp0 && !fp1 && fq0 && !fq1 && mv_far(mp0, mq0, ylim)) return 1;
    if (fp0 && !fp1 && !fq0 && fq1 && mv_far(mp0, mq1, ylim)) return 1;
    if (!fp0 && fp1 && fq0 && !fq1 && mv_far(mp1, mq0, ylim)) return 1;
    if (!fp0 && fp1 && !fq0 && fq1 && mv_far(mp1, mq1, ylim)) return 1;
    if (fp0 && fp1 && p0 != p1 && fq0 && fq1 && ((q0 == p0 && q1 == p1) || (q0 == p1 && q1 == p0))) {
        if (q0 == p0 && (mv_far(mp0, mq0, ylim) || mv_far(mp1, mq1, ylim))) return 1;
        if (q0 == p1 && (mv_far(mp1, mq0, ylim) || mv_far(mp0, mq1, ylim))) return 1;
    }
    if (fp0 && fp1 && p0 == p1 && fq0 && fq1 && q0 == q1 && q0 == p0) {
        /* Q3: && binds tighter than || here (Deblock.cpp:1289-1299) */
        if (mv_far(mp0, mq0, ylim) || (mv_far(mp1, mq1, ylim) && mv_far(mp0, mq1, ylim)) || mv_far(mp1, mq0, ylim)) return 1;
    }
    return 0;
}

/* 8.7.2.1 (Deblock.cpp:994-1310) for luma positions (xp, yp) in MB p and (xq, yq) in MB q
 * of a non-MBAFF frame. */
static int bs_luma(const MbInfo *P, const MbInfo *Q, int mb_edge, int xp, int yp, int xq, int yq) {
    int pintra = is_intra_cls(P->cls), qintra = is_intra_cls(Q->cls);
    if (mb_edge && (pintra || qintra)) return 4;
    if (pintra || qintra) return 3;
    if (bs_nonzero(P, Q, xp, yp, xq, yq)) return 2;
    return bs_motion(P, Q, xp, yp, xq, yq, 4);
}

/* 8.7.2.1 in an MBAFF frame (Deblock.cpp:994-1310): mixedModeEdgeFlag when p and q lie in
 * a frame and a field MB; bS 4 for MB edges between frame MBs and for vertical MB edges,
 * 3 for the other intra edges, 1 for mixed edges without levels. */
static int bs_mbaff(const Ctx *X, int ap, int aq, int xp, int yp, int xq, int yq, int vertical) {
    const MbInfo *P = &X->info[ap], *Q = &X->info[aq];
    int intra = is_intra_cls(P->cls) || is_intra_cls(Q->cls);
    int mixed = ap != aq && P->field != Q->field;
    if (ap != aq && intra && ((!P->field && !Q->field) || vertical)) return 4;
    if (intra && (!mixed || !vertical)) return 3;
    if (bs_nonzero(P, Q, xp, yp, xq, yq)) return 2;
    if (mixed) return 1;
    return bs_motion(P, Q, xp, yp, xq, yq, Q->field ? 2 : 4);
}

/* 8.7.2.2 - 8.7.2.4 (Deblock.cpp:1314-1522) on p[0..3] / q[0..3] */
void refcpu_filter_line(uint8_t p[4], uint8_t q[4], int bS, int chroma, int qp_avg, int offA, int offB) {
    int indexA = CLIP3(0, 51, qp_avg + offA), indexB = CLIP3(0, 51, qp_avg + offB);
    int alpha = ALPHA[indexA], beta = BETA[indexB];
    int p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3], q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    if (!(bS != 0 && iabs(p0 - q0) < alpha && iabs(p1 - p0) < beta && iabs(q1 - q0) < beta)) return;
    int ap = iabs(p2 - p0), aq = iabs(q2 - q0);
    if (bS < 4) {
        int tc0 = TC0[bS - 1][indexA];
        int tc = chroma ? tc0 + 1 : tc0 + (ap < beta) + (aq < beta);
        int delta = CLIP3(-tc, tc, (((q0 - p0) << 2) + (p1 - q1) + 4) >> 3);
        p[0] = (uint8_t)CLIP1(p0 + delta);
        q[0] = (uint8_t)CLIP1(q0 - delta);
        if (!chroma && ap < beta) p[1] = (uint8_t)(p1 + CLIP3(-tc0, tc0, (p2 + ((p0 + q0 + 1) >> 1) - (p1 << 1)) >> 1));
        if (!chroma && aq < beta) q[1] = (uint8_t)(q1 + CLIP3(-tc0, tc0, (q2 + ((p0 + q0 + 1) >> 1) - (q1 << 1)) >> 1));
        return;
    }
    int strong = iabs(p0 - q0) < ((alpha >> 2) + 2);
    if (!chroma && ap < beta && strong) {
        p[0] = (uint8_t)((p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3);
        p[1] = (uint8_t)((p2 + p1 + p0 + q0 + 2) >> 2);
        p[2] = (uint8_t)((2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3);
    } else {
        p[0] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
    }
    if (!chroma && aq < beta && strong) {
        q[0] = (uint8_t)((p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3);
        q[1] = (uint8_t)((p0 + q0 + q1 + q2 + 2) >> 2);
        q[2] = (uint8_t)((2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3);
    } else {
        q[0] = (uint8_t)((2 * q1 + q0 + p1 + 2) >> 2);
    }
}

/* One edge of MB (mbx, mby): plane c (0 luma, 1 Cb, 2 Cr), vertical or horizontal, at
 * offset e inside the MB (Deblock.cpp:639-840).  own: an MB edge whose p-MB is the current
 * one (Q7): the bS of an internal edge between the MB's blocks at 15 and 0, the MB's qP on
 * both sides, the neighbour's samples. */
static void filter_edge(Ctx *X, int mbx, int mby, int c, int vertical, int e, int own) {
    int n = c ? 8 : 16;
    uint8_t *plane = c == 0 ? X->y : (c == 1 ? X->cb : X->cr);
    int stride = c ? X->wc : X->wl;
    const MbInfo *Q = &X->info[mby * X->W + mbx];
    const MbInfo *P = Q;
    int mb_edge = e == 0 && !own;
    if (mb_edge) P = vertical ? &X->info[mby * X->W + mbx - 1] : &X->info[(mby - 1) * X->W + mbx];
    int scale = c ? 2 : 1;
    int qpp, qpq;
    if (c == 0) {
        qpp = P->cls == H264G_MB_IPCM ? 0 : P->qp;
        qpq = Q->cls == H264G_MB_IPCM ? 0 : Q->qp;
    } else {
        /* Filtering_process_for_block_edges passes iCbCr (0 = Cb) where the callee expects
         * isChromaCb (1 = Cb), so the edge QPs of Cb use the Cr offset and vice versa
         * (Deblock.cpp:813-815, 918-941; H264PictureBase.cpp:4816-4829). */
        int off = c == 1 ? X->d->second_chroma_qp_index_offset : X->d->chroma_qp_index_offset;
        qpp = refcpu_chroma_qp(P->cls == H264G_MB_IPCM ? 0 : P->qp, off);
        qpq = refcpu_chroma_qp(Q->cls == H264G_MB_IPCM ? 0 : Q->qp, off);
    }
    int qavg = (qpp + qpq + 1) >> 1;
    for (int k = 0; k < n; k++) {
        /* sample positions inside the MBs, then mapped to luma for bS (Deblock.cpp:871-879) */
        int xq = vertical ? e : k, yq = vertical ? k : e;
        int xp = vertical ? (e ? e - 1 : n - 1) : k, yp = vertical ? k : (e ? e - 1 : n - 1);
        int bS = bs_luma(P, Q, mb_edge, xp * scale, yp * scale, xq * scale, yq * scale);
        uint8_t pp[4], qq[4];
        int gx = mbx * n + xq, gy = mby * n + yq;
        for (int i = 0; i < 4; i++) {
            if (vertical) { qq[i] = plane[gy * stride + gx + i]; pp[i] = plane[gy * stride + gx - i - 1]; }
            else { qq[i] = plane[(gy + i) * stride + gx]; pp[i] = plane[(gy - i - 1) * stride + gx]; }
        }
        refcpu_filter_line(pp, qq, bS, c != 0, qavg, X->d->filter_offset_a, X->d->filter_offset_b);
        for (int i = 0; i < 3; i++) {
            if (vertical) { plane[gy * stride + gx + i] = qq[i]; plane[gy * stride + gx - i - 1] = pp[i]; }
            else { plane[(gy + i) * stride + gx] = qq[i]; plane[(gy - i - 1) * stride + gx] = pp[i]; }
        }
    }
}

/* One edge pass of MB addr in an MBAFF frame (Deblock.cpp:639-990): plane c; the lines
 * E[k] = (e, k) (vertical) or (k, e) (horizontal); sample rows dy = 1 + fieldMode apart
 * from the MB's position (yE % 2 selects the bottom field lines of a frame MB's top edge
 * over a field pair); mbAddrN: the p-MB of an MB edge, alternating with the line's
 * parity when leftMbEdge (frame MB beside a field pair).  The p0 position keeps the
 * current MB's line numbering (Q6). */
static void filter_edge_mbaff(Ctx *X, int addr, int c, int vertical, int e, int fieldMode, int leftMbEdge, int mbAddrN) {
    const int chroma = c != 0, nE = chroma ? 8 : 16, dy = 1 + fieldMode, s = chroma ? 2 : 1;
    uint8_t *plane = c == 0 ? X->y : (c == 1 ? X->cb : X->cr);
    const int stride = chroma ? X->wc : X->wl;
    int xI, yI;
    mb_luma_pos(X, addr, &xI, &yI);
    const int xP = chroma ? xI / 2 : xI, yP = chroma ? (yI + 1) / 2 : yI;
    const MbInfo *Q = &X->info[addr];
    /* Cb edges take the Cr offset and vice versa (Q16, see filter_edge) */
    const int off = c == 1 ? X->d->second_chroma_qp_index_offset : X->d->chroma_qp_index_offset;
    for (int k = 0; k < nE; k++) {
        int xE = vertical ? e : k, yE = vertical ? k : e;
        uint8_t pp[4], qq[4];
        int qo[4], po[4];
        for (int i = 0; i < 4; i++) {
            if (vertical) {
                qo[i] = (yP + dy * yE) * stride + xP + xE + i;
                po[i] = (yP + dy * yE) * stride + xP + xE - i - 1;
            } else {
                qo[i] = (yP + dy * (yE + i) - yE % 2) * stride + xP + xE;
                po[i] = (yP + dy * (yE - i - 1) - yE % 2) * stride + xP + xE;
            }
            qq[i] = plane[qo[i]];
            pp[i] = plane[po[i]];
        }
        int ap = addr, xp, yp, xq, yq;
        if (vertical) {
            xp = xE - 1;
            if (xp < 0) {
                if (mbAddrN >= 0) ap = leftMbEdge ? mbAddrN + yE % 2 : mbAddrN;
                xp += nE;
            }
            yp = yE; xq = xE; yq = yE;
        } else {
            xp = xE;
            yp = (yE - 1) - yE % 2;
            if (yp < 0) {
                if (mbAddrN >= 0) ap = mbAddrN;
                yp += nE;
            }
            xq = xE; yq = yE - yE % 2;
        }
        const MbInfo *P = &X->info[ap];
        int bS = bs_mbaff(X, ap, addr, xp * s, yp * s, xq * s, yq * s, vertical);
        int qpp = P->cls == H264G_MB_IPCM ? 0 : P->qp, qpq = Q->cls == H264G_MB_IPCM ? 0 : Q->qp;
        if (chroma) { qpp = refcpu_chroma_qp(qpp, off); qpq = refcpu_chroma_qp(qpq, off); }
        refcpu_filter_line(pp, qq, bS, chroma, (qpp + qpq + 1) >> 1, X->d->filter_offset_a, X->d->filter_offset_b);
        for (int i = 0; i < 3; i++) {
            plane[qo[i]] = qq[i];
            plane[po[i]] = pp[i];
        }
    }
}

/* 8.7 for MBAFF frames, MBs in address order (Deblock.cpp:76-635): left / top edges
 * when the pair has a left / upper neighbour (a frame MB's top edge inside its pair
 * always); a frame MB under a field pair filters its top edge twice in field mode, top
 * field lines against the pair's top MB, bottom field lines against its bottom MB
 * (chroma: the bottom MB both times, Q5). */
static void deblock_mbaff(Ctx *X) {
    const int idc = X->d->disable_deblocking_filter_idc;
    if (idc == 1) return;
    for (int a = 0; a < X->W * X->H; a++) {
        int xW, yW;
        const int mbAddrA = mbaff_nb(X, a, -1, 0, 16, 16, &xW, &yW), mbAddrB = mbaff_nb(X, a, 0, -1, 16, 16, &xW, &yW);
        const int field = mb_field(X, a), top = a % 2 == 0, row0 = (a >> 1) < X->W;
        const int t8 = X->info[a].t8;
        const int left = (a >> 1) % X->W != 0 && !(idc == 2 && mbAddrA < 0);
        const int upper = !(row0 && (field || top)) && !(idc == 2 && mbAddrB < 0);
        const int leftMix = a >= 2 && !field && mb_field(X, a - 2);
        const int topMix = top && a >= 2 * X->W && !field && mb_field(X, a - 2 * X->W + 1);
        if (left) filter_edge_mbaff(X, a, 0, 1, 0, field, leftMix, mbAddrA);
        if (!t8) filter_edge_mbaff(X, a, 0, 1, 4, field, 0, mbAddrA);
        filter_edge_mbaff(X, a, 0, 1, 8, field, 0, mbAddrA);
        if (!t8) filter_edge_mbaff(X, a, 0, 1, 12, field, 0, mbAddrA);
        if (upper && topMix) {
            filter_edge_mbaff(X, a, 0, 0, 0, 1, 0, mbAddrB - 1);
            filter_edge_mbaff(X, a, 0, 0, 1, 1, 0, mbAddrB);
        } else if (upper) {
            filter_edge_mbaff(X, a, 0, 0, 0, field, 0, mbAddrB);
        }
        if (!t8) filter_edge_mbaff(X, a, 0, 0, 4, field, 0, mbAddrB);
        filter_edge_mbaff(X, a, 0, 0, 8, field, 0, mbAddrB);
        if (!t8) filter_edge_mbaff(X, a, 0, 0, 12, field, 0, mbAddrB);
        for (int c = 1; c <= 2; c++) {
            if (left) filter_edge_mbaff(X, a, c, 1, 0, field, leftMix, mbAddrA);
            filter_edge_mbaff(X, a, c, 1, 4, field, 0, mbAddrA);
            if (upper && topMix) {
                filter_edge_mbaff(X, a, c, 0, 0, 1, 0, mbAddrB);
                filter_edge_mbaff(X, a, c, 0, 1, 1, 0, mbAddrB);
            } else if (upper) {
                filter_edge_mbaff(X, a, c, 0, 0, field, 0, mbAddrB);
            }
            filter_edge_mbaff(X, a, c, 0, 4, field, 0, mbAddrB);
        }
    }
}

/* 8.7 in MB raster order (Deblock.cpp:76-635), up to the first MB of no slice (Q22) */
static void deblock(Ctx *X) {
    const int idc = X->d->disable_deblocking_filter_idc;
    if (idc == 1) return;
    for (int mby = 0; mby < X->H; mby++)
        for (int mbx = 0; mbx < X->W; mbx++) {
            const int a = mby * X->W + mbx;
            const MbInfo *I = &X->info[a];
            const uint32_t sn = X->mbs[a].slice_num;
            if (sn == H264G_SLICE_NONE) return;
            const int otherL = mbx > 0 && X->mbs[a - 1].slice_num != sn;
            const int otherT = mby > 0 && X->mbs[a - X->W].slice_num != sn;
            int left = mbx > 0 && !(idc == 2 && otherL), topf = mby > 0 && !(idc == 2 && otherT);
            const int ownL = idc == 0 && otherL, ownT = idc == 0 && otherT;  /* Q7 */
            if (left) filter_edge(X, mbx, mby, 0, 1, 0, ownL);
            if (!I->t8) filter_edge(X, mbx, mby, 0, 1, 4, 0);
            filter_edge(X, mbx, mby, 0, 1, 8, 0);
            if (!I->t8) filter_edge(X, mbx, mby, 0, 1, 12, 0);
            if (topf) filter_edge(X, mbx, mby, 0, 0, 0, ownT);
            if (!I->t8) filter_edge(X, mbx, mby, 0, 0, 4, 0);
            filter_edge(X, mbx, mby, 0, 0, 8, 0);
            if (!I->t8) filter_edge(X, mbx, mby, 0, 0, 12, 0);
            for (int c = 1; c <= 2; c++) {
                if (left) filter_edge(X, mbx, mby, c, 1, 0, ownL);
                filter_edge(X, mbx, mby, c, 1, 4, 0);
                if (topf) filter_edge(X, mbx, mby, c, 0, 0, ownT);
                filter_edge(X, mbx, mby, c, 0, 4, 0);
            }
        }
}

/* ------------------------------------------------------ picture */
int refcpu_decode_picture(const h264g_pic_desc *d, const h264g_mb *mbs, int W, int H, const int16_t *levels,
                          uint8_t *const *planes, int stage) {
    if (d->structure != H264G_FRAME || (d->mbaff && (H % 2 || d->constrained_intra_pred_flag))) return H264G_E_ARG;
    Ctx X;
    X.d = d; X.mbs = mbs; X.W = W; X.H = H;
    X.wl = W * 16; X.hl = H * 16; X.wc = W * 8; X.hc = H * 8;
    X.planes = planes;
    X.mbaff = d->mbaff ? 1 : 0;
    if (X.mbaff && d->scaling_matrix) return H264G_E_ARG;  /* field MBs would map the lists with the field scan */
    X.y = planes[d->dst_slot];
    X.cb = X.y + X.wl * X.hl;
    X.cr = X.cb + X.wc * X.hc;
    X.info = (MbInfo *)calloc((size_t)W * H, sizeof(MbInfo));
    Levels *L = (Levels *)malloc(sizeof(Levels));
    if (!X.info || !L) { free(X.info); free(L); return H264G_E_NOMEM; }
    int rc = 0;
    for (int a = 0; a < W * H && !rc; a++) {
        const h264g_mb *mb = &mbs[a];
        MbInfo *info = &X.info[a];
        const int field = mb_field(&X, a);
        if (X.mbaff && (mbs[a ^ 1].flags & H264G_MBF_FIELD) != (mb->flags & H264G_MBF_FIELD)) { rc = H264G_E_ARG; break; }
        if (!X.mbaff && (mb->flags & H264G_MBF_FIELD)) { rc = H264G_E_ARG; break; }
        parse_levels(mb, levels, L);
        info->cls = mb->mb_type;
        info->qp = mb->qp_y;
        info->field = field;
        info->t8 = (mb->flags & H264G_MBF_T8x8) && mb->mb_type != H264G_MB_IPCM ? 1 : 0;
        for (int b = 0; b < 16; b++) {
            int nz = 0;
            for (int k = 0; k < 16; k++) nz |= L->l4[b][k] != 0;
            info->nz4 |= (uint16_t)(nz << b);
        }
        for (int q = 0; q < 4; q++) {
            int nz = 0;
            for (int k = 0; k < 64; k++) nz |= L->l8[q][k] != 0;
            info->nz8 |= (uint8_t)(nz << q);
        }
        for (int l = 0; l < 2; l++)
            for (int b = 0; b < 16; b++) info->refid[l][b] = -1;
        int x0, y0, pitch, cx0, cy0, cpitch;
        mb_plane_origin(&X, a, 16, X.wl, &x0, &y0, &pitch);
        mb_plane_origin(&X, a, 8, X.wc, &cx0, &cy0, &cpitch);
        uint8_t *oy = X.y + y0 * X.wl + x0, *ocb = X.cb + cy0 * X.wc + cx0, *ocr = X.cr + cy0 * X.wc + cx0;
        if (mb->mb_type == H264G_MB_IPCM) {
            /* RefIdxLX = 0 from the zeroed MB state (see file header); samples in the MB's
             * rows (H264PictureBase.cpp:2449-2499) */
            for (int l = 0; l < 2; l++)
                for (int b = 0; b < 16; b++) info->refid[l][b] = d->num_ref_idx_active[l] > 0 ? d->ref_slot[l][0] : -1;
            for (int i = 0; i < 256; i++) oy[(i / 16) * pitch + i % 16] = (uint8_t)L->pcm[i];
            for (int i = 0; i < 64; i++) {
                ocb[(i / 8) * cpitch + i % 8] = (uint8_t)L->pcm[256 + i];
                ocr[(i / 8) * cpitch + i % 8] = (uint8_t)L->pcm[320 + i];
            }
        } else if (is_intra_cls(mb->mb_type)) {
            rc = intra_mb(&X, a, mb, L);
        } else {
            rc = inter_mb(&X, a, mb, info);
            if (!rc && mb->mb_type != H264G_MB_P_SKIP) {
                /* residual added to the prediction (H264InterPrediction.cpp:22-407); scaling
                 * matrices: the inter luma lists, chroma included (Q13) */
                int qp = mb->qp_y;
                const uint8_t *w4 = d->scaling_matrix ? d->weight_scale4x4[3] : NULL;
                const uint8_t *w8 = d->scaling_matrix ? d->weight_scale8x8[1] : NULL;
                if (info->t8) {
                    for (int q = 0; q < 4; q++) {
                        int32_t r[64];
                        residual8x8_w(L->l8[q], qp, field, w8, r);
                        add_residual(oy, pitch, (q % 2) * 8, (q / 2) * 8, r, 8);
                    }
                } else {
                    for (int b = 0; b < 16; b++) {
                        int32_t r[16];
                        int bx, by;
                        blk_xy(b, &bx, &by);
                        luma4x4_residual(L, b, qp, field, w4, r);
                        add_residual(oy, pitch, bx, by, r, 4);
                    }
                }
                for (int c = 0; c < 2; c++) {
                    int32_t rmb[64];
                    int off = c == 0 ? d->chroma_qp_index_offset : d->second_chroma_qp_index_offset;
                    chroma_residual(L, c, refcpu_chroma_qp(qp, off), field, w4, rmb);
                    add_residual(c == 0 ? ocb : ocr, cpitch, 0, 0, rmb, 8);
                }
            }
        }
    }
    if (!rc && stage >= 2) {
        if (X.mbaff) deblock_mbaff(&X);
        else deblock(&X);
    }
    free(X.info);
    free(L);
    (void)imin; (void)imax;
    return rc;
}

/* ------------------------------------------------ output conversion */
/* H264PictureBase.cpp:440-466 (and the bottom-up variant at :469-495): BT.601 integer
 * coefficients over 1000, C division (truncation toward zero), clipped to 0..255;
 * chroma sampled at (x / 2, y / 2). */
void refcpu_yuv420p_to_bgr24(int width, int height, const uint8_t *yuv, uint8_t *bgr, int stride, int flip) {
    const uint8_t *u = yuv + (size_t)width * height, *v = u + (size_t)width * height / 4;
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            const int Y = yuv[(size_t)y * width + x] - 16;
            const int U = u[(size_t)(y / 2) * (width / 2) + x / 2] - 128, V = v[(size_t)(y / 2) * (width / 2) + x / 2] - 128;
            const int b = (1164 * Y + 2018 * U) / 1000, g = (1164 * Y - 813 * V - 391 * U) / 1000, r = (1164 * Y + 1596 * V) / 1000;
            uint8_t *o = bgr + (size_t)(flip ? height - 1 - y : y) * stride + (size_t)x * 3;
            o[0] = (uint8_t)CLIP3(0, 255, b);
            o[1] = (uint8_t)CLIP3(0, 255, g);
            o[2] = (uint8_t)CLIP3(0, 255, r);
        }
}
