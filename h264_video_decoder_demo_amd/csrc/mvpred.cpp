// mvpred.cpp -- host motion-vector derivation for frame pictures (8.4.1).
//
// A serial host parser reads mvd_l0 / mvd_l1 per (sub-)partition; the final motion
// vector is mvp + mvd, with mvp predicted from the already decoded neighbouring
// partitions (8.4.1.3), and P_Skip MBs carry no mvd at all (8.4.1.1).  The reference
// derives both inside its reconstruction (P_Skip: H264InterPrediction.cpp:685-738;
// prediction: Derivation_process_for_luma_motion_vector_prediction and its
// neighbouring-partition process, H264InterPrediction.cpp:1660-2001), which this
// library's device path does not have: the h264g_mb record carries final vectors.
// h264g_mvfield restates that derivation on the host so that a parser can turn its
// mvds into records.  Frame pictures only (no MBAFF field scaling); B_Skip / B_Direct
// (8.4.1.2) are not derived, as the record cannot carry them.
//
// The field keeps, per 4x4 luma block of the picture, the reference index and vector
// of each list and whether the block's partition has been decoded.  A neighbouring
// partition (6.4.11.7) in another MB is available when its block has been decoded and
// its MB lies in the current MB's slice (h264g_mb.slice_num, 6.4.8): MBs are derived in
// decoding (raster) order, so an MB outside the picture, the MB to the right and the MBs
// below are never decoded.
//
// Inside the current MB the reference departs from 6.4.11.7: its test that marks a
// not yet decoded partition as not available is commented out
// (H264InterPrediction.cpp:3164-3169), so such a partition counts as available, and its
// motion is read from the MB's arrays as parsing left them: PredFlagLX 0, hence refIdx
// -1 and mv 0 (H264InterPrediction.cpp:1862-1873, 593-603).  Only neighbour C can fall
// there (in P_8x8 / B_8x8, a sub-partition of quadrant 0 or 2 whose C lies in
// quadrant 1 or 3), and it is then not replaced by D.  The field reproduces that.
//
// The field also predicts the luma intra modes of I_NxN MBs (8.3.1.1 / 8.3.2.1), which a
// parser reads as prev_intra*_pred_mode_flag / rem_intra*_pred_mode: per MB it keeps the
// prediction class and the sixteen 4x4 modes (an I8x8 MB's four modes, each on its four
// 4x4 blocks), as the reference keeps m_mb_pred_mode and Intra4x4PredMode /
// Intra8x8PredMode (H264PictureBase.cpp:773-1058).
//
// Spatial direct prediction (8.4.1.2.2, H264InterPrediction.cpp:1259-1432) of B_Skip,
// B_Direct_16x16 and B_Direct_8x8 (records with H264G_MBF_DIRECT bits): reference
// indices MinPositive over the MB's neighbours A, B and C (or D), the 16x16 median
// prediction per list, and zero vectors where a list's index is 0 and the quadrant's
// co-located block in RefPicList1[0] (set by h264g_mvfield_colocated; its corner 4x4
// block under direct_8x8_inference_flag 1) has index 0 and a vector within +-1.  As in
// the reference (H264InterPrediction.cpp:1226-1249) an intra co-located MB has index -1
// but an I_PCM one, which is not "intra" there, reads its zeroed motion: index 0, vector
// 0 (quirk Q20).  Temporal direct prediction is not derived: the reference's copy
// divides by zero (quirk Q19).
#include "h264g.h"
#include "h264g_internal.h"

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

namespace {

struct Blk {
    int8_t ref[2];     // -1: the list is not used (or the MB is intra)
    uint8_t decoded;
    int16_t mv[2][2];
};

struct Nb {  // one neighbouring partition for one list (8.4.1.3.2)
    bool avail;
    int ref;
    int mv[2];
};

inline int median3(int a, int b, int c) { return a + b + c - std::min(a, std::min(b, c)) - std::max(a, std::max(b, c)); }

}  // namespace

// Prediction class of an MB for intra-mode prediction (the reference's m_mb_pred_mode).
// CLS_INTER: Pred_L0 / Pred_L1 / BiPred, which dcPredModePredictedFlag tests under
// constrained intra prediction; P_8x8 / B_8x8 have Pred_NA (H264MacroBlock.cpp:99-100,
// 142) and are CLS_OTHER like Intra_16x16 and I_PCM.
enum : uint8_t { CLS_NONE = 0, CLS_INTER, CLS_I4, CLS_I8, CLS_OTHER };

// luma4x4BlkIdx of the 4x4 block holding luma sample (x, y) of an MB (6.4.13.1).
inline int blk4_at(int x, int y) { return 4 * ((y >> 3) * 2 + (x >> 3)) + ((y >> 2) & 1) * 2 + ((x >> 2) & 1); }

struct h264g_mvfield {
    int W = 0, H = 0;  // in MBs
    std::vector<Blk> b;  // (4W) x (4H) blocks, raster
    std::vector<uint8_t> cls;    // per MB: CLS_*, CLS_NONE until derived in this picture
    std::vector<uint8_t> imode;  // per MB: 16 modes by luma4x4BlkIdx (I4x4 / I8x8 MBs)
    std::vector<uint8_t> colzero;  // per MB: bit q = colZeroFlag of quadrant q (spatial direct)
    std::vector<uint32_t> slice;   // per MB: slice_num of the derived MBs
    uint32_t cur_slice = 0;        // slice_num of the MB being derived
    bool have_col = false;

    Blk *at(int bx, int by) { return &b[(size_t)by * W * 4 + bx]; }

    // intraMxMPredModeN of the neighbour holding luma sample (x, y) of MB (mx, my), with x
    // or y possibly -1 (the MB left / above); `cur` holds the current MB's modes derived
    // so far.  blk8: the 8x8 derivation, whose I4x4 neighbours give their 4x4 block n of
    // the 8x8 block holding the sample (n = 1 left, 2 above; H264PictureBase.cpp:1010-1037).
    // dc: dcPredModePredictedFlag (unavailable neighbour, or inter neighbour under
    // constrained intra prediction).
    int intra_nb(int mx, int my, int x, int y, bool blk8, int n, uint8_t cur_cls, const uint8_t *cur, bool cip,
                 bool &dc) {
        int ax = mx, ay = my;
        if (x < 0) { ax--; x += 16; }
        if (y < 0) { ay--; y += 16; }
        uint8_t c;
        const uint8_t *modes;
        if (ax == mx && ay == my) {
            c = cur_cls;
            modes = cur;
        } else {
            if (ax < 0 || ay < 0 || ax >= W || cls[(size_t)ay * W + ax] == CLS_NONE ||
                slice[(size_t)ay * W + ax] != cur_slice) {
                dc = true;
                return 2;
            }
            c = cls[(size_t)ay * W + ax];
            modes = &imode[((size_t)ay * W + ax) * 16];
        }
        if (cip && c == CLS_INTER) dc = true;
        if (c != CLS_I4 && c != CLS_I8) return 2;
        const int k = blk4_at(x, y);
        if (blk8 && c == CLS_I4) return modes[(k & ~3) + n];
        return modes[k];  // I8x8 modes sit on every 4x4 block of their 8x8 block
    }
    // Partition covering 4x4 block (mx*4 + x4, my*4 + y4), for list l.
    Nb nb(int mx, int my, int x4, int y4, int l) {
        Nb n{false, -1, {0, 0}};
        const int bx = mx * 4 + x4, by = my * 4 + y4;
        if (bx < 0 || by < 0 || bx >= W * 4 || by >= H * 4) return n;
        const Blk *k = at(bx, by);
        const bool in_cur = x4 >= 0 && x4 < 4 && y4 >= 0 && y4 < 4;
        if (!k->decoded) {
            n.avail = in_cur;  // current MB: available, no motion
            return n;
        }
        if (!in_cur && slice[(size_t)(by >> 2) * W + (bx >> 2)] != cur_slice) return n;  // another slice
        n.avail = true;
        n.ref = k->ref[l];
        if (n.ref >= 0) {
            n.mv[0] = k->mv[l][0];
            n.mv[1] = k->mv[l][1];
        }
        return n;
    }

    // 8.4.1.3: mvp of the partition at 4x4 position (x4, y4), predPartWidth w4 (4x4
    // units), for list l and reference index ref.  dir: 1 = 16x8 partition, 2 = 8x16
    // partition (with part its index), 0 otherwise.
    void mvp(int mx, int my, int x4, int y4, int w4, int l, int ref, int dir, int part, int out[2]) {
        const Nb A = nb(mx, my, x4 - 1, y4, l), B = nb(mx, my, x4, y4 - 1, l);
        Nb C = nb(mx, my, x4 + w4, y4 - 1, l);
        if (!C.avail) C = nb(mx, my, x4 - 1, y4 - 1, l);  // D replaces an unavailable C
        const Nb *dirn = nullptr;  // directional prediction of 16x8 / 8x16 (8.4.1.3)
        if (dir == 1) dirn = part == 0 ? &B : &A;
        else if (dir == 2) dirn = part == 0 ? &A : &C;
        if (dirn && dirn->ref == ref) {
            out[0] = dirn->mv[0];
            out[1] = dirn->mv[1];
            return;
        }
        // 8.4.1.3.1 median
        Nb a = A, bb = B, c = C;
        if (!bb.avail && !c.avail && a.avail) bb = c = a;
        const int n = (a.ref == ref) + (bb.ref == ref) + (c.ref == ref);
        if (n == 1) {
            const Nb &m = a.ref == ref ? a : (bb.ref == ref ? bb : c);
            out[0] = m.mv[0];
            out[1] = m.mv[1];
            return;
        }
        for (int k = 0; k < 2; k++) out[k] = median3(a.mv[k], bb.mv[k], c.mv[k]);
    }

    // Writes the motion of the w4 x h4 blocks at (x4, y4) of MB (mx, my) and marks them decoded.
    void put(int mx, int my, int x4, int y4, int w4, int h4, const int ref[2], const int mv[2][2], h264g_mb *mb) {
        for (int y = y4; y < y4 + h4; y++)
            for (int x = x4; x < x4 + w4; x++) {
                Blk *k = at(mx * 4 + x, my * 4 + y);
                k->decoded = 1;
                for (int l = 0; l < 2; l++) {
                    k->ref[l] = (int8_t)ref[l];
                    for (int c = 0; c < 2; c++) {
                        k->mv[l][c] = (int16_t)(ref[l] >= 0 ? mv[l][c] : 0);
                        if (mb) mb->mv[l][y * 4 + x][c] = k->mv[l][c];
                    }
                }
            }
    }
};

extern "C" {

int32_t h264g_mvfield_create(int32_t width_mbs, int32_t height_mbs, h264g_mvfield **out) {
    if (!out || width_mbs < 1 || height_mbs < 1 || width_mbs > 1024 || height_mbs > 1024)
        return h264g_set_error(H264G_E_ARG, "mvfield %d x %d MBs", width_mbs, height_mbs);
    h264g_mvfield *f = new (std::nothrow) h264g_mvfield;
    if (!f) return h264g_set_error(H264G_E_NOMEM, "mvfield");
    f->W = width_mbs;
    f->H = height_mbs;
    try {
        f->b.assign((size_t)width_mbs * height_mbs * 16, Blk{{-1, -1}, 0, {{0, 0}, {0, 0}}});
        f->cls.assign((size_t)width_mbs * height_mbs, CLS_NONE);
        f->imode.assign((size_t)width_mbs * height_mbs * 16, 2);
        f->colzero.assign((size_t)width_mbs * height_mbs, 0);
        f->slice.assign((size_t)width_mbs * height_mbs, 0);
    } catch (...) {
        delete f;
        return h264g_set_error(H264G_E_NOMEM, "mvfield");
    }
    *out = f;
    return 0;
}

void h264g_mvfield_destroy(h264g_mvfield *f) { delete f; }

int32_t h264g_mvfield_reset(h264g_mvfield *f) {
    if (!f) return h264g_set_error(H264G_E_ARG, "mvfield_reset: null field");
    std::fill(f->b.begin(), f->b.end(), Blk{{-1, -1}, 0, {{0, 0}, {0, 0}}});
    std::fill(f->cls.begin(), f->cls.end(), (uint8_t)CLS_NONE);
    return 0;
}

int32_t h264g_mvfield_colocated(h264g_mvfield *f, const h264g_mb *col_mbs, int32_t col_short_term) {
    if (!f) return h264g_set_error(H264G_E_ARG, "mvfield_colocated: null field");
    f->have_col = col_mbs != nullptr;
    if (!col_mbs) return 0;
    static const int corner[4] = {0, 3, 12, 15};  // luma4x4BlkIdx 0, 5, 10, 15 in raster order
    for (size_t a = 0; a < f->colzero.size(); a++) {
        const h264g_mb &c = col_mbs[a];
        uint8_t z = 0;
        for (int q = 0; q < 4 && col_short_term; q++) {
            bool cz;
            if (c.mb_type == H264G_MB_I4x4 || c.mb_type == H264G_MB_I8x8 || c.mb_type == H264G_MB_I16x16) {
                cz = false;  // refIdxCol -1
            } else if (c.mb_type == H264G_MB_IPCM) {
                cz = true;   // Q20: refIdxCol 0, mvCol 0
            } else {         // predFlagL0Col ? L0 : L1 of the partition holding the corner block
                const int lc = c.ref_idx[0][q] >= 0 ? 0 : 1;
                const int16_t *v = c.mv[lc][corner[q]];
                cz = c.ref_idx[lc][q] == 0 && v[0] >= -1 && v[0] <= 1 && v[1] >= -1 && v[1] <= 1;
            }
            if (cz) z |= (uint8_t)(1u << q);
        }
        f->colzero[a] = z;
    }
    return 0;
}

int32_t h264g_mvfield_intra_modes(h264g_mvfield *f, uint32_t mb_addr, h264g_mb *mb, const uint8_t *syntax,
                                  int32_t constrained_intra_pred) {
    if (!f || !mb || !syntax) return h264g_set_error(H264G_E_ARG, "mvfield_intra_modes: null argument");
    if (mb_addr >= (uint32_t)(f->W * f->H)) return h264g_set_error(H264G_E_ARG, "mvfield_intra_modes: MB %u", mb_addr);
    const bool i8 = mb->mb_type == H264G_MB_I8x8;
    if (!i8 && mb->mb_type != H264G_MB_I4x4)
        return h264g_set_error(H264G_E_ARG, "mvfield_intra_modes: MB %u type %d is not I_NxN", mb_addr, mb->mb_type);
    const int mx = (int)(mb_addr % (uint32_t)f->W), my = (int)(mb_addr / (uint32_t)f->W);
    const bool cip = constrained_intra_pred != 0;
    f->cur_slice = mb->slice_num;
    uint8_t cur[16];
    std::memset(cur, 2, sizeof(cur));
    const int nblk = i8 ? 4 : 16;
    for (int k = 0; k < nblk; k++) {
        // the block's top-left sample (6.4.3 / 6.4.5), its neighbours A (left) and B (above)
        const int x = i8 ? (k & 1) * 8 : ((k >> 2) & 1) * 8 + (k & 1) * 4;
        const int y = i8 ? (k >> 1) * 8 : ((k >> 3) & 1) * 8 + ((k >> 1) & 1) * 4;
        bool dc = false;
        const int a = f->intra_nb(mx, my, x - 1, y, i8, 1, i8 ? CLS_I8 : CLS_I4, cur, cip, dc);
        const int b = f->intra_nb(mx, my, x, y - 1, i8, 2, i8 ? CLS_I8 : CLS_I4, cur, cip, dc);
        const int pred = dc ? 2 : std::min(a, b);
        const int prev = syntax[2 * k], rem = syntax[2 * k + 1];
        if (!prev && rem > 7) return h264g_set_error(H264G_E_ARG, "mvfield_intra_modes: MB %u rem %d", mb_addr, rem);
        const int mode = prev ? pred : (rem < pred ? rem : rem + 1);
        mb->intra_pred_modes[k] = (uint8_t)mode;
        if (i8) std::memset(&cur[4 * k], mode, 4);
        else cur[k] = (uint8_t)mode;
    }
    return 0;
}

int32_t h264g_mvfield_derive(h264g_mvfield *f, uint32_t mb_addr, h264g_mb *mb, const int16_t *mvd) {
    if (!f || !mb) return h264g_set_error(H264G_E_ARG, "mvfield_derive: null argument");
    if (mb_addr >= (uint32_t)(f->W * f->H)) return h264g_set_error(H264G_E_ARG, "mvfield_derive: MB %u", mb_addr);
    const int mx = (int)(mb_addr % (uint32_t)f->W), my = (int)(mb_addr / (uint32_t)f->W);
    f->cur_slice = mb->slice_num;
    f->slice[mb_addr] = mb->slice_num;
    for (int y = 0; y < 4; y++)  // a partition of this MB is available only once derived
        for (int x = 0; x < 4; x++) f->at(mx * 4 + x, my * 4 + y)->decoded = 0;
    {  // the MB's class and intra modes, for the intra-mode prediction of later MBs
        const int t = mb->mb_type;
        uint8_t *im = &f->imode[(size_t)mb_addr * 16];
        const bool split = (t == H264G_MB_P || t == H264G_MB_B) && mb->part_shape == H264G_PART_8x8;
        f->cls[mb_addr] = t == H264G_MB_I4x4 ? CLS_I4 : t == H264G_MB_I8x8 ? CLS_I8
                        : (t == H264G_MB_I16x16 || t == H264G_MB_IPCM || split) ? CLS_OTHER : CLS_INTER;
        for (int k = 0; k < 16; k++)
            im[k] = t == H264G_MB_I4x4 ? mb->intra_pred_modes[k] : t == H264G_MB_I8x8 ? mb->intra_pred_modes[k >> 2] : 2;
    }
    const int none[2] = {-1, -1}, zero[2][2] = {{0, 0}, {0, 0}};
    switch (mb->mb_type) {
    case H264G_MB_I4x4:
    case H264G_MB_I8x8:
    case H264G_MB_I16x16:
    case H264G_MB_IPCM:  // neighbours see an intra MB as refIdx -1, mv 0 (8.4.1.3.2)
        f->put(mx, my, 0, 0, 4, 4, none, zero, nullptr);
        return 0;
    case H264G_MB_P_SKIP: {  // 8.4.1.1 (H264InterPrediction.cpp:685-738)
        const Nb A = f->nb(mx, my, -1, 0, 0), B = f->nb(mx, my, 0, -1, 0);
        int mv[2][2] = {{0, 0}, {0, 0}};
        if (A.avail && B.avail && !(A.ref == 0 && A.mv[0] == 0 && A.mv[1] == 0) &&
            !(B.ref == 0 && B.mv[0] == 0 && B.mv[1] == 0))
            f->mvp(mx, my, 0, 0, 4, 0, 0, 0, 0, mv[0]);
        for (int q = 0; q < 4; q++) {
            mb->ref_idx[0][q] = 0;
            mb->ref_idx[1][q] = -1;
        }
        const int ref[2] = {0, -1};
        f->put(mx, my, 0, 0, 4, 4, ref, mv, mb);
        return 0;
    }
    case H264G_MB_P:
    case H264G_MB_B:
        break;
    default:
        return h264g_set_error(H264G_E_ARG, "mvfield_derive: MB %u type %d", mb_addr, mb->mb_type);
    }
    const uint8_t direct = mb->mb_type == H264G_MB_B ? (uint8_t)(mb->flags & H264G_MBF_DIRECT_ALL) >> 4 : 0;
    if (direct && (mb->part_shape != H264G_PART_8x8 || !f->have_col))
        return h264g_set_error(H264G_E_ARG, "mvfield_derive: MB %u: direct needs part_shape 8x8 and a co-located picture",
                               mb_addr);
    int dref[2] = {-1, -1}, dmv[2][2] = {{0, 0}, {0, 0}};
    bool dzero = false;
    if (direct) {  // 8.4.1.2.2, once for the MB: its neighbours lie outside it
        for (int l = 0; l < 2; l++) {
            const Nb A = f->nb(mx, my, -1, 0, l), B = f->nb(mx, my, 0, -1, l);
            Nb C = f->nb(mx, my, 4, -1, l);
            if (!C.avail) C = f->nb(mx, my, -1, -1, l);
            auto min_pos = [](int a, int b) { return a >= 0 && b >= 0 ? std::min(a, b) : std::max(a, b); };
            dref[l] = min_pos(A.ref, min_pos(B.ref, C.ref));
        }
        dzero = dref[0] < 0 && dref[1] < 0;
        if (dzero) dref[0] = dref[1] = 0;
        for (int l = 0; l < 2 && !dzero; l++)
            if (dref[l] >= 0) f->mvp(mx, my, 0, 0, 4, l, dref[l], 0, 0, dmv[l]);
    }
    if (!mvd && direct != 0xF) return h264g_set_error(H264G_E_ARG, "mvfield_derive: MB %u needs mvd", mb_addr);
    const int shape = mb->part_shape;
    if (shape > H264G_PART_8x8) return h264g_set_error(H264G_E_ARG, "mvfield_derive: MB %u shape %d", mb_addr, shape);
    const int np = shape == H264G_PART_16x16 ? 1 : (shape == H264G_PART_8x8 ? 4 : 2);
    for (int p = 0; p < np; p++) {
        // partition p (mbPartIdx) in 4x4 units, then its sub-partitions in order
        int px = 0, py = 0, pw = 4, ph = 4, ns = 1, sw = 4, sh = 4;
        if (shape == H264G_PART_16x8) { py = 2 * p; ph = 2; }
        else if (shape == H264G_PART_8x16) { px = 2 * p; pw = 2; }
        else if (shape == H264G_PART_8x8) {
            px = 2 * (p & 1); py = 2 * (p >> 1); pw = ph = 2;
            const int ss = mb->sub_part_shape[p];
            if (ss > H264G_SUB_4x4) return h264g_set_error(H264G_E_ARG, "mvfield_derive: MB %u sub shape %d", mb_addr, ss);
            ns = ss == H264G_SUB_8x8 ? 1 : (ss == H264G_SUB_4x4 ? 4 : 2);
        }
        sw = pw;  // (sub-)partition size: predPartWidth is the sub-partition's width for P_8x8 / B_8x8
        sh = ph;
        if (shape == H264G_PART_8x8) {
            const int ss = mb->sub_part_shape[p];
            sw = (ss == H264G_SUB_4x8 || ss == H264G_SUB_4x4) ? 1 : 2;
            sh = (ss == H264G_SUB_8x4 || ss == H264G_SUB_4x4) ? 1 : 2;
        }
        const int q = (py >> 1) * 2 + (px >> 1);  // 8x8 quadrant of the partition's first block
        if (direct & (1u << p)) {  // a direct quadrant: indices and vectors from the MB's prediction
            const bool cz = (f->colzero[mb_addr] >> p) & 1;
            int mv[2][2] = {{0, 0}, {0, 0}};
            for (int l = 0; l < 2; l++) {
                mb->ref_idx[l][q] = (int8_t)dref[l];
                if (dref[l] >= 0 && !dzero && !(dref[l] == 0 && cz)) { mv[l][0] = dmv[l][0]; mv[l][1] = dmv[l][1]; }
            }
            f->put(mx, my, px, py, 2, 2, dref, mv, mb);
            continue;
        }
        const int ref[2] = {mb->ref_idx[0][q], mb->ref_idx[1][q]};
        if (ref[0] < 0 && ref[1] < 0)
            return h264g_set_error(H264G_E_ARG, "mvfield_derive: MB %u partition %d uses no list", mb_addr, p);
        const int dir = shape == H264G_PART_16x8 ? 1 : (shape == H264G_PART_8x16 ? 2 : 0);
        for (int s = 0; s < ns; s++) {
            int sx = px, sy = py;  // top-left 4x4 block of sub-partition s
            if (shape == H264G_PART_8x8) {
                if (sw == 1 && sh == 1) { sx += s & 1; sy += s >> 1; }
                else if (sw == 1) sx += s;
                else if (sh == 1) sy += s;
            }
            const int blk = sy * 4 + sx;
            int mv[2][2] = {{0, 0}, {0, 0}};
            for (int l = 0; l < 2; l++) {
                if (ref[l] < 0) continue;
                int pred[2];
                f->mvp(mx, my, sx, sy, sw, l, ref[l], dir, p, pred);
                for (int c = 0; c < 2; c++) mv[l][c] = pred[c] + mvd[(l * 16 + blk) * 2 + c];
            }
            f->put(mx, my, sx, sy, sw, sh, ref, mv, mb);
        }
    }
    return 0;
}

}  // extern "C"
