// gen.cpp -- deterministic synthetic H.264 syntax for the benchmark workloads.
//
// The generator produces, per picture in decoding order, exactly what a serial
// host parser would hand to the engine: the picture descriptor (slice header /
// PPS / reference-list state) and one h264g_mb record per macroblock plus the
// coefficient-level stream.  Syntax is valid for the reference decoder:
//   * intra prediction modes are drawn only among the modes whose neighbouring
//     samples are available (the reference writes nothing for the others,
//     H264PictureBase.cpp:1176-1186),
//   * QPY follows mb_qp_delta semantics (7-37): MBs without mb_qp_delta inherit
//     QPY,PREV (H264MacroBlock.cpp:1098-1124, 1194-1198),
//   * P_Skip MVs are the 8.4.1.1 prediction (H264InterPrediction.cpp:685-738),
//   * reference lists follow the default initialisation of 8.2.4.2 for frames.
#include "h264g.h"
#include "h264g_internal.h"

#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

// ---------------------------------------------------------------- random
struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; i++) s[i] = splitmix(x);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {  // xoshiro256**
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform integer in [lo, hi]
    int32_t uni(int32_t lo, int32_t hi) {
        uint64_t span = (uint64_t)((int64_t)hi - lo + 1);
        return (int32_t)(lo + (int64_t)(next() % span));
    }
    bool permille(int32_t p) { return (int32_t)(next() % 1000) < p; }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline int clip3(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }

// ------------------------------------------------------------- GOP layout
struct PicPlan {
    int slice_type;
    int display;     // display index
    int is_ref;
    int idr;
    int frame_num;   // decoding-order index inside the GOP (PicNum for frames)
};

std::vector<PicPlan> plan_gop(int n_frames, int b_frames, int gop) {
    // Anchors (I/P) at display indices 0, b+1, 2(b+1), ... and the last frame;
    // each anchor is decoded before the B pictures that precede it in display order.
    std::vector<PicPlan> out;
    if (gop <= 0) gop = n_frames;
    for (int g0 = 0; g0 < n_frames; g0 += gop) {
        int g1 = std::min(n_frames, g0 + gop);
        int fn = 0;
        int prev_anchor = g0;
        out.push_back({H264G_SLICE_I, g0, 1, 1, fn++});
        while (prev_anchor < g1 - 1) {
            int a = std::min(prev_anchor + b_frames + 1, g1 - 1);
            out.push_back({H264G_SLICE_P, a, 1, 0, fn++});
            for (int d = prev_anchor + 1; d < a; d++) out.push_back({H264G_SLICE_B, d, 0, 0, fn});
            prev_anchor = a;
        }
    }
    return out;
}

struct Picture {
    h264g_pic_desc desc;
    std::vector<h264g_mb> mbs;
    std::vector<int16_t> levels;
};

// Per-4x4 motion state of the picture being generated (for P_Skip prediction).
struct Motion {
    int16_t mv[2][2];
    int8_t ref[2];
    uint8_t intra;
};

}  // namespace

struct h264g_gen {
    h264g_gen_params p;
    int dpb_slots;
    std::vector<Picture> pics;
    uint64_t digest;
};

namespace {

// Appends one MB's levels (masks first, then values) from dense scan-order blocks.
struct LevelWriter {
    std::vector<uint16_t> masks;
    std::vector<int16_t> values;
    void block(const int16_t *coef, int n, int first) {
        // n <= 16 entries per mask word; 64-entry blocks call this four times.
        uint16_t m = 0;
        for (int k = first; k < n; k++)
            if (coef[k]) { m |= (uint16_t)(1u << k); values.push_back(coef[k]); }
        masks.push_back(m);
    }
    void flush(std::vector<int16_t> &out) {
        for (uint16_t m : masks) out.push_back((int16_t)m);
        out.insert(out.end(), values.begin(), values.end());
        masks.clear(); values.clear();
    }
};

class PictureGen {
public:
    // col: the records of RefPicList1[0] (B pictures with direct MBs), else null
    PictureGen(const h264g_gen_params &p, Rng &rng, h264g_pic_desc &desc, Picture &pic, const h264g_mb *col = nullptr)
        : P(p), R(rng), D(desc), pic(pic), W(p.width_mbs), H(p.height_mbs), col(col) {
        motion.assign((size_t)W * 4 * H * 4, Motion());
    }

    void run() {
        int slice_qp = R.uni(P.qp_lo, P.qp_hi);
        qp_prev = slice_qp;
        pic.mbs.resize((size_t)W * H);
        if (P.mbaff) {
            run_mbaff();
            return;
        }
        for (int my = 0; my < H; my++)
            for (int mx = 0; mx < W; mx++) gen_mb(mx, my);
    }

private:
    const h264g_gen_params &P;
    Rng &R;
    h264g_pic_desc &D;
    Picture &pic;
    int W, H;
    const h264g_mb *col;
    int qp_prev = 26;
    std::vector<Motion> motion;    // per 4x4 block: picture raster (frames), or [MB address][raster
                                   // block] (MBAFF frames)
    std::vector<uint8_t> fld;      // MBAFF: mb_field_decoding_flag per MB address

    Motion &mo(int bx, int by) { return motion[(size_t)by * W * 4 + bx]; }
    bool mb_exists(int mx, int my) const { return mx >= 0 && my >= 0 && mx < W && my < H; }

    int16_t draw_level() {
        if (R.permille(P.zero_permille)) return 0;
        double b = P.laplace_b_x100 / 100.0;
        double u = R.unit();
        if (u < 1e-300) u = 1e-300;
        int mag = 1 + (int)std::floor(-b * std::log(u));
        if (mag > 64) mag = 64;
        return (int16_t)((R.next() & 1) ? -mag : mag);
    }

    int draw_mv_comp() {
        if (P.mv_mode == 0) return R.uni(-P.mv_range, P.mv_range);
        double u1 = R.unit(), u2 = R.unit();
        if (u1 < 1e-300) u1 = 1e-300;
        double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        int v = (int)std::lround(g * P.mv_range);
        return clip3(P.mv_clamp_lo, P.mv_clamp_hi, v);
    }

    // --- intra mode choice among modes whose neighbours are available ---------
    int pick_nxn_mode(bool left, bool top, bool topleft) {
        int cand[9], n = 0;
        cand[n++] = 2;                                   // DC
        if (top) { cand[n++] = 0; cand[n++] = 3; cand[n++] = 7; }
        if (left) { cand[n++] = 1; cand[n++] = 8; }
        if (top && left && topleft) { cand[n++] = 4; cand[n++] = 5; cand[n++] = 6; }
        return cand[R.uni(0, n - 1)];
    }
    int pick_16_mode(bool left, bool top, bool topleft) {
        int cand[4], n = 0;
        cand[n++] = 2;
        if (top) cand[n++] = 0;
        if (left) cand[n++] = 1;
        if (top && left && topleft) cand[n++] = 3;
        return cand[R.uni(0, n - 1)];
    }
    int pick_chroma_mode(bool left, bool top, bool topleft) {
        int cand[4], n = 0;
        cand[n++] = 0;
        if (left) cand[n++] = 1;
        if (top) cand[n++] = 2;
        if (top && left && topleft) cand[n++] = 3;
        return cand[R.uni(0, n - 1)];
    }

    // --- 8.4.1.1 P_Skip motion vector (non-MBAFF frames) -----------------------
    // Neighbour motion is read per 4x4 block, which equals MvL0[mbPartIdxN][subMbPartIdxN]
    // and RefIdxL0[mbPartIdxN] of the covering partition (H264InterPrediction.cpp:1757-2001).
    void neighbour(int bx, int by, bool &avail, int &ref, int mv[2]) {
        mv[0] = mv[1] = 0;
        ref = -1;
        avail = bx >= 0 && by >= 0 && bx < W * 4 && by < H * 4;
        if (!avail) return;
        const Motion &m = mo(bx, by);
        if (m.intra || m.ref[0] < 0) return;
        ref = m.ref[0];
        mv[0] = m.mv[0][0];
        mv[1] = m.mv[0][1];
    }

    void pskip_mv(int mx, int my, int out[2]) {
        int x0 = mx * 4, y0 = my * 4;
        bool aA, aB, aC;
        int rA, rB, rC, mA[2], mB[2], mC[2];
        neighbour(x0 - 1, y0, aA, rA, mA);
        neighbour(x0, y0 - 1, aB, rB, mB);
        // C at (16, -1): the macroblock above-right, else D at (-1, -1).
        bool c_exists = mb_exists(mx + 1, my - 1);
        if (c_exists) neighbour(x0 + 4, y0 - 1, aC, rC, mC);
        else neighbour(x0 - 1, y0 - 1, aC, rC, mC);
        if (!aA || !aB || (rA == 0 && mA[0] == 0 && mA[1] == 0) ||
            (rB == 0 && mB[0] == 0 && mB[1] == 0)) {
            out[0] = out[1] = 0;
            return;
        }
        // median prediction with refIdx = 0 (8.4.1.3.1)
        if (!aB && !aC && aA) {
            mB[0] = mC[0] = mA[0]; mB[1] = mC[1] = mA[1];
            rB = rC = rA;
        }
        int cnt = (rA == 0) + (rB == 0) + (rC == 0);
        if (cnt == 1) {
            const int *m = rA == 0 ? mA : (rB == 0 ? mB : mC);
            out[0] = m[0]; out[1] = m[1];
            return;
        }
        for (int c = 0; c < 2; c++) {
            int a = mA[c], b = mB[c], cc = mC[c];
            out[c] = a + b + cc - std::min(a, std::min(b, cc)) - std::max(a, std::max(b, cc));
        }
    }

    // --- one macroblock -------------------------------------------------------
    static void init_mb(h264g_mb &mb) {
        std::memset(&mb, 0, sizeof(mb));
        for (int l = 0; l < 2; l++)
            for (int q = 0; q < 4; q++) mb.ref_idx[l][q] = -1;
    }

    void gen_mb(int mx, int my) {
        h264g_mb &mb = pic.mbs[(size_t)my * W + mx];
        init_mb(mb);
        if (P.slices > 1) {
            const int run = (W * H + P.slices - 1) / P.slices;
            mb.slice_num = (uint32_t)((my * W + mx) / run);
        }
        // neighbours for intra prediction: inside the picture and, under constrained intra
        // prediction, not an MB whose prediction mode the reference counts as inter
        // (Pred_L0 / Pred_L1 / BiPred: P_Skip and the P / B MBs not split into 8x8; P_8x8
        // and B_8x8 have Pred_NA and stay available)
        auto avail = [&](int x, int y) {
            if (!mb_exists(x, y)) return false;
            const h264g_mb &n = pic.mbs[(size_t)y * W + x];
            if (n.slice_num != mb.slice_num) return false;  // another slice
            const bool hides = n.mb_type == H264G_MB_P_SKIP ||
                               ((n.mb_type == H264G_MB_P || n.mb_type == H264G_MB_B) && n.part_shape != H264G_PART_8x8);
            return !P.constrained_intra_pred || !hides;
        };
        const bool A = avail(mx - 1, my), B = avail(mx, my - 1), Dn = avail(mx - 1, my - 1);
        const int cls = choose_class(mb);
        gen_body(mb, cls, A, B, Dn, false, [&](int mv[2]) { pskip_mv(mx, my, mv); },
                 [&](h264g_mb &m) { direct_spatial(mx, my, m); });
        if (cls <= H264G_MB_IPCM) mark_intra(mx, my);
        else store_motion(mx, my, mb);
    }

    // Macroblock class (and partition shape of P MBs of the config-1 mix).
    int choose_class(h264g_mb &mb) {
        const int st = D.slice_type;
        int cls;
        if (st == H264G_SLICE_I) {
            cls = R.permille(500) ? H264G_MB_I16x16 : (P.t8x8 && R.permille(500) ? H264G_MB_I8x8 : H264G_MB_I4x4);
        } else if (P.p_pattern && st == H264G_SLICE_P) {
            int u = R.uni(0, 999);
            cls = u < 100 ? H264G_MB_P_SKIP : H264G_MB_P;
            if (cls == H264G_MB_P) mb.part_shape = (u < 400) ? H264G_PART_8x8 : H264G_PART_16x16;
        } else {
            if (R.permille(P.intra_in_inter_permille)) {
                int u = R.uni(0, 2);
                cls = u == 0 ? H264G_MB_I16x16 : (u == 1 ? H264G_MB_I4x4 : (P.t8x8 ? H264G_MB_I8x8 : H264G_MB_I4x4));
            } else if (st == H264G_SLICE_P && R.permille(P.skip_permille)) {
                cls = H264G_MB_P_SKIP;
            } else if (st == H264G_SLICE_B && P.direct_permille > 0 && R.permille(P.direct_permille)) {
                // B_Skip, B_Direct_16x16, or B_8x8 with B_Direct_8x8 sub-macroblocks
                cls = H264G_MB_B;
                mb.part_shape = H264G_PART_8x8;
                const int u = R.uni(0, 2);
                mb.flags |= u == 2 ? (uint8_t)(R.uni(1, 14) << 4) : (uint8_t)H264G_MBF_DIRECT_ALL;
                if (u == 0) mb.flags |= H264G_MBF_SKIP;
            } else {
                cls = st == H264G_SLICE_P ? H264G_MB_P : H264G_MB_B;
                int u = R.uni(0, 99);
                mb.part_shape = u < 40 ? H264G_PART_16x16 : (u < 55 ? H264G_PART_16x8 : (u < 70 ? H264G_PART_8x16 : H264G_PART_8x8));
            }
        }
        if (P.ipcm_permille && cls != H264G_MB_P_SKIP && R.permille(P.ipcm_permille)) {
            cls = H264G_MB_IPCM;
            mb.flags &= (uint8_t)~(H264G_MBF_SKIP | H264G_MBF_DIRECT_ALL);  // no longer a direct B MB
        }
        mb.mb_type = (uint8_t)cls;
        return cls;
    }

    // Everything of an MB after its class: prediction syntax, levels, QP.  A / B / Dn:
    // availability of the neighbouring MBs for intra prediction; field: a field MB of an
    // MBAFF frame (its reference indices address the field list, twice as long);
    // pskip(mv): the P_Skip motion vector.
    template <class PSkip, class Direct>
    void gen_body(h264g_mb &mb, int cls, bool A, bool B, bool Dn, bool field, PSkip pskip, Direct direct) {
        bool intra = cls <= H264G_MB_IPCM;
        LevelWriter lw;
        mb.level_offset = (uint32_t)pic.levels.size();

        if (cls == H264G_MB_IPCM) {
            for (int i = 0; i < 384; i++) pic.levels.push_back((int16_t)R.uni(0, 255));
            mb.qp_y = (uint8_t)qp_prev;  // mb_qp_delta not present: QPY = QPY,PREV
            return;
        }

        // ---- prediction syntax
        if (cls == H264G_MB_I4x4 || cls == H264G_MB_I8x8) {
            int n = cls == H264G_MB_I4x4 ? 4 : 2;   // blocks per row
            int step = 4 / n;                       // in 4x4 units
            for (int by = 0; by < n; by++)
                for (int bx = 0; bx < n; bx++) {
                    bool left = bx > 0 || A, top = by > 0 || B;
                    bool tl = (bx > 0 && by > 0) || (bx == 0 && by > 0 && A) || (bx > 0 && by == 0 && B) || (bx == 0 && by == 0 && Dn);
                    int mode = pick_nxn_mode(left, top, tl);
                    int blk = cls == H264G_MB_I4x4 ? blk4x4_index(bx * step, by * step) : (by * 2 + bx);
                    mb.intra_pred_modes[blk] = (uint8_t)mode;
                }
            if (cls == H264G_MB_I8x8) mb.flags |= H264G_MBF_T8x8;
        } else if (cls == H264G_MB_I16x16) {
            mb.intra16x16_pred_mode = (uint8_t)pick_16_mode(A, B, Dn);
        }
        if (intra) mb.intra_chroma_pred_mode = (uint8_t)pick_chroma_mode(A, B, Dn);

        if (cls == H264G_MB_P_SKIP) {
            int mv[2];
            pskip(mv);
            for (int b = 0; b < 16; b++) { mb.mv[0][b][0] = (int16_t)mv[0]; mb.mv[0][b][1] = (int16_t)mv[1]; }
            for (int q = 0; q < 4; q++) mb.ref_idx[0][q] = 0;
            mb.qp_y = (uint8_t)qp_prev;
            return;
        }
        if (cls == H264G_MB_P || cls == H264G_MB_B) gen_inter(mb, field);
        if (mb.flags & H264G_MBF_DIRECT_ALL) direct(mb);
        if (mb.flags & H264G_MBF_SKIP) {  // B_Skip: no residual, mb_qp_delta absent
            mb.qp_y = (uint8_t)qp_prev;
            return;
        }

        // ---- transform size for inter MBs (eligible: no sub-8x8 partitions)
        bool inter = cls == H264G_MB_P || cls == H264G_MB_B;
        if (inter && P.t8x8) {
            bool eligible = true;
            if (mb.part_shape == H264G_PART_8x8)
                for (int q = 0; q < 4; q++) if (mb.sub_part_shape[q] != H264G_SUB_8x8) eligible = false;
            if (eligible && R.permille(500)) mb.flags |= H264G_MBF_T8x8;
        }
        bool t8 = (mb.flags & H264G_MBF_T8x8) != 0;

        // ---- residual levels (dense, scan order) and CBP derived from them
        int16_t dc16[16] = {0};
        int16_t luma[16][16];      // 4x4 blocks by luma4x4BlkIdx (or I16 AC at 1..15)
        int16_t luma8[4][64];
        int16_t cdc[2][4], cac[2][4][16];
        std::memset(luma, 0, sizeof(luma));
        std::memset(luma8, 0, sizeof(luma8));
        std::memset(cac, 0, sizeof(cac));
        if (cls == H264G_MB_I16x16) for (int k = 0; k < 16; k++) dc16[k] = draw_level();
        int cbp_luma = 0;
        for (int q = 0; q < 4; q++) {
            bool nz = false;
            if (t8) {
                for (int k = 0; k < 64; k++) { luma8[q][k] = draw_level(); nz |= luma8[q][k] != 0; }
            } else {
                for (int b = 0; b < 4; b++)
                    for (int k = (cls == H264G_MB_I16x16 ? 1 : 0); k < 16; k++) {
                        luma[q * 4 + b][k] = draw_level();
                        nz |= luma[q * 4 + b][k] != 0;
                    }
            }
            if (nz) cbp_luma |= 1 << q;
        }
        if (cls == H264G_MB_I16x16 && cbp_luma) cbp_luma = 15;  // I16x16 signals 0 or 15
        bool cdc_nz = false, cac_nz = false;
        for (int c = 0; c < 2; c++) {
            for (int k = 0; k < 4; k++) { cdc[c][k] = draw_level(); cdc_nz |= cdc[c][k] != 0; }
            for (int b = 0; b < 4; b++)
                for (int k = 1; k < 16; k++) { cac[c][b][k] = draw_level(); cac_nz |= cac[c][b][k] != 0; }
        }
        int cbp_chroma = cac_nz ? 2 : (cdc_nz ? 1 : 0);
        if (cbp_chroma < 2) std::memset(cac, 0, sizeof(cac));
        if (inter && cbp_luma == 0) mb.flags &= (uint8_t)~H264G_MBF_T8x8;  // not signalled without luma CBP
        t8 = (mb.flags & H264G_MBF_T8x8) != 0;
        mb.cbp = (uint8_t)(cbp_luma | (cbp_chroma << 4));

        // ---- QP: mb_qp_delta present only with residual or for I16x16
        bool has_delta = cbp_luma || cbp_chroma || cls == H264G_MB_I16x16;
        if (has_delta) qp_prev = R.uni(P.qp_lo, P.qp_hi);
        mb.qp_y = (uint8_t)qp_prev;

        // ---- emit levels
        if (cls == H264G_MB_I16x16) lw.block(dc16, 16, 0);
        for (int q = 0; q < 4; q++) {
            if (!(cbp_luma & (1 << q))) continue;
            if (t8) {
                for (int w = 0; w < 4; w++) {
                    int16_t part[16];
                    std::memcpy(part, &luma8[q][w * 16], sizeof(part));
                    lw.block(part, 16, 0);
                }
            } else {
                for (int b = 0; b < 4; b++) lw.block(luma[q * 4 + b], 16, cls == H264G_MB_I16x16 ? 1 : 0);
            }
        }
        if (cbp_chroma >= 1) for (int c = 0; c < 2; c++) lw.block(cdc[c], 4, 0);
        if (cbp_chroma == 2)
            for (int c = 0; c < 2; c++)
                for (int b = 0; b < 4; b++) lw.block(cac[c][b], 16, 1);
        lw.flush(pic.levels);
    }

    static int blk4x4_index(int x4, int y4) {  // luma4x4BlkIdx of the 4x4 at (x4, y4)
        return 8 * (y4 / 2) + 4 * (x4 / 2) + 2 * (y4 % 2) + (x4 % 2);
    }

    void mark_intra(int mx, int my) {
        for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) {
                Motion &m = mo(mx * 4 + x, my * 4 + y);
                m = Motion();
                m.intra = 1;
                m.ref[0] = m.ref[1] = -1;
            }
    }

    void store_motion(int mx, int my, const h264g_mb &mb) {
        for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) {
                Motion &m = mo(mx * 4 + x, my * 4 + y);
                int q = (y / 2) * 2 + (x / 2);
                m.intra = 0;
                for (int l = 0; l < 2; l++) {
                    m.ref[l] = mb.ref_idx[l][q];
                    m.mv[l][0] = mb.mv[l][y * 4 + x][0];
                    m.mv[l][1] = mb.mv[l][y * 4 + x][1];
                }
            }
    }

    // --- 8.4.1.2.2 spatial direct prediction (frames, direct_8x8_inference_flag 1) --
    // (H264InterPrediction.cpp:1259-1432.)  Reference indices: MinPositive over the
    // neighbouring partitions A, B and C (D when C is not available) of the MB as one
    // 16-wide partition; both negative: 0 / 0 with zero vectors.  Vectors: the 16x16
    // median prediction of each list, zero where the list's index is 0 and the quadrant's
    // co-located block (the corner 4x4 block of the same quadrant in RefPicList1[0],
    // 8.4.1.2.1) has reference index 0 and a vector within +-1.  As in the reference
    // (H264InterPrediction.cpp:1226-1249) an intra co-located MB has index -1, except
    // I_PCM, which is not intra there and reads its zeroed motion: index 0, vector 0.
    void nb_list(int bx, int by, int l, bool &avail, int &ref, int mv[2]) {
        mv[0] = mv[1] = 0;
        ref = -1;
        avail = bx >= 0 && by >= 0 && bx < W * 4 && by < H * 4;
        if (!avail) return;
        const Motion &m = mo(bx, by);
        if (m.intra || m.ref[l] < 0) return;
        ref = m.ref[l];
        mv[0] = m.mv[l][0];
        mv[1] = m.mv[l][1];
    }

    void direct_spatial(int mx, int my, h264g_mb &mb) {
        const int x0 = mx * 4, y0 = my * 4;
        const bool c_exists = mb_exists(mx + 1, my - 1);
        bool av[3][2];
        int rf[3][2], mv[3][2][2];
        for (int l = 0; l < 2; l++) {
            nb_list(x0 - 1, y0, l, av[0][l], rf[0][l], mv[0][l]);
            nb_list(x0, y0 - 1, l, av[1][l], rf[1][l], mv[1][l]);
            if (c_exists) nb_list(x0 + 4, y0 - 1, l, av[2][l], rf[2][l], mv[2][l]);
            else nb_list(x0 - 1, y0 - 1, l, av[2][l], rf[2][l], mv[2][l]);
        }
        auto min_pos = [](int a, int b) { return a >= 0 && b >= 0 ? std::min(a, b) : std::max(a, b); };
        int ref[2], mvp[2][2] = {{0, 0}, {0, 0}};
        for (int l = 0; l < 2; l++) ref[l] = min_pos(rf[0][l], min_pos(rf[1][l], rf[2][l]));
        const bool zero = ref[0] < 0 && ref[1] < 0;
        if (zero) ref[0] = ref[1] = 0;
        for (int l = 0; l < 2 && !zero; l++) {
            if (ref[l] < 0) continue;
            int r[3], m[3][2];
            for (int k = 0; k < 3; k++) { r[k] = rf[k][l]; m[k][0] = mv[k][l][0]; m[k][1] = mv[k][l][1]; }
            if (!av[1][l] && !av[2][l] && av[0][l]) {  // 8.4.1.3.1: B and C take A
                for (int k = 1; k < 3; k++) { r[k] = r[0]; m[k][0] = m[0][0]; m[k][1] = m[0][1]; }
            }
            const int n = (r[0] == ref[l]) + (r[1] == ref[l]) + (r[2] == ref[l]);
            for (int c = 0; c < 2; c++) {
                if (n == 1) mvp[l][c] = r[0] == ref[l] ? m[0][c] : (r[1] == ref[l] ? m[1][c] : m[2][c]);
                else mvp[l][c] = m[0][c] + m[1][c] + m[2][c] - std::min(m[0][c], std::min(m[1][c], m[2][c])) -
                                 std::max(m[0][c], std::max(m[1][c], m[2][c]));
            }
        }
        const h264g_mb &cm = col[(size_t)my * W + mx];
        static const int corner[4] = {0, 3, 12, 15};
        for (int q = 0; q < 4; q++) {
            if (!(mb.flags & H264G_MBF_DIRECT(q))) continue;
            bool colzero;
            if (cm.mb_type == H264G_MB_I4x4 || cm.mb_type == H264G_MB_I8x8 || cm.mb_type == H264G_MB_I16x16) {
                colzero = false;
            } else if (cm.mb_type == H264G_MB_IPCM) {
                colzero = true;
            } else {
                const int lc = cm.ref_idx[0][q] >= 0 ? 0 : 1;
                const int16_t *v = cm.mv[lc][corner[q]];
                colzero = cm.ref_idx[lc][q] == 0 && v[0] >= -1 && v[0] <= 1 && v[1] >= -1 && v[1] <= 1;
            }
            const int qx = (q & 1) * 2, qy = (q >> 1) * 2;
            for (int l = 0; l < 2; l++) {
                mb.ref_idx[l][q] = (int8_t)ref[l];
                const bool z = zero || ref[l] < 0 || (ref[l] == 0 && colzero);
                for (int y = qy; y < qy + 2; y++)
                    for (int x = qx; x < qx + 2; x++)
                        for (int c = 0; c < 2; c++) mb.mv[l][y * 4 + x][c] = ref[l] < 0 ? 0 : (int16_t)(z ? 0 : mvp[l][c]);
            }
        }
    }

    // --- MBAFF frames ----------------------------------------------------------
    // MB address 2 p + b is the top (b = 0) or bottom MB of pair p (pairs in raster order);
    // both MBs of a pair share mb_field_decoding_flag.  Motion is kept per MB address.
    Motion &mo_at(int addr, int blk) { return motion[(size_t)addr * 16 + blk]; }

    // 6.4.12.2 (Table 6-4) for a luma location (xN, yN) relative to MB addr: the MB
    // covering it (-1: not available) and the location (xW, yW) inside that MB.  Pairs A,
    // B, C, D of 6.4.10 are available when inside the picture (one slice per picture).
    int mbaff_nb(int addr, int xN, int yN, int &xW, int &yW) const {
        const int p = addr >> 1, px = p % W, top = (addr & 1) == 0;
        const int pA = px > 0 ? p - 1 : -1, pB = p >= W ? p - W : -1;
        const int pC = (p >= W && px < W - 1) ? p - W + 1 : -1, pD = (p >= W && px > 0) ? p - W - 1 : -1;
        const bool cur_field = fld[addr] != 0;
        auto field_pair = [&](int q) { return fld[2 * q] != 0; };
        int n = -1, yM = 0;
        if (xN > 15 && yN >= 0) return -1;
        if (yN > 15) return -1;
        if (yN >= 0 && xN >= 0) {
            n = addr;
            yM = yN;
        } else if (xN < 0 && yN >= 0) {  // left pair
            if (pA < 0) return -1;
            const bool af = field_pair(pA);
            if (!cur_field) {
                if (!af) { n = 2 * pA + (top ? 0 : 1); yM = yN; }
                else { n = 2 * pA + (yN & 1); yM = (yN + (top ? 0 : 16)) >> 1; }
            } else if (!af) {
                const int r = 2 * yN + (top ? 0 : 1);  // row of the frame pair
                n = 2 * pA + (r >= 16 ? 1 : 0);
                yM = r & 15;
            } else {
                n = 2 * pA + (top ? 0 : 1);
                yM = yN;
            }
        } else {  // yN < 0: above, above-left, above-right
            const int q = xN < 0 ? -1 : (xN < 16 ? 0 : 1);
            if (!cur_field && !top) {  // bottom frame MB: inside the pair, or the left pair
                if (q == 1) return -1;
                if (q == 0) { n = addr - 1; yM = yN; }
                else {
                    if (pA < 0) return -1;
                    if (!field_pair(pA)) { n = 2 * pA; yM = yN; }
                    else { n = 2 * pA + 1; yM = (yN + 16) >> 1; }
                }
            } else {
                const int pn = q < 0 ? pD : (q == 0 ? pB : pC);
                if (pn < 0) return -1;
                if (!cur_field || !top) { n = 2 * pn + 1; yM = yN; }          // frame top MB, bottom field MB
                else if (!field_pair(pn)) { n = 2 * pn + 1; yM = 2 * yN; }   // top field MB over a frame pair
                else { n = 2 * pn; yM = yN; }
            }
        }
        xW = (xN + 16) % 16;
        yW = (yM + 16) % 16;
        return n;
    }

    // Motion data of the partition covering luma location (xN, yN) of MB addr for list 0,
    // scaled to the current MB's frame / field units (8.4.1.3.2 with 6.4.11.7,
    // H264InterPrediction.cpp:1758-2001): field MB over a frame neighbour halves the
    // vertical component (C division) and doubles refIdx, the other way round doubles
    // the component and halves refIdx.
    void neighbour_mbaff(int addr, int xN, int yN, bool &avail, int &ref, int mv[2]) {
        mv[0] = mv[1] = 0;
        ref = -1;
        int xW = 0, yW = 0;
        const int n = mbaff_nb(addr, xN, yN, xW, yW);
        avail = n >= 0;
        if (!avail) return;
        const Motion &m = mo_at(n, (yW >> 2) * 4 + (xW >> 2));
        if (m.intra || m.ref[0] < 0) return;
        ref = m.ref[0];
        mv[0] = m.mv[0][0];
        mv[1] = m.mv[0][1];
        if (fld[addr] && !fld[n]) { mv[1] = mv[1] / 2; ref = ref * 2; }
        else if (!fld[addr] && fld[n]) { mv[1] = mv[1] * 2; ref = ref / 2; }
    }

    // 8.4.1.1 P_Skip motion vector in an MBAFF frame (H264InterPrediction.cpp:685-738).
    void pskip_mv_mbaff(int addr, int out[2]) {
        bool aA, aB, aC;
        int rA, rB, rC, mA[2], mB[2], mC[2];
        neighbour_mbaff(addr, -1, 0, aA, rA, mA);
        neighbour_mbaff(addr, 0, -1, aB, rB, mB);
        neighbour_mbaff(addr, 16, -1, aC, rC, mC);
        if (!aC) neighbour_mbaff(addr, -1, -1, aC, rC, mC);
        if (!aA || !aB || (rA == 0 && mA[0] == 0 && mA[1] == 0) || (rB == 0 && mB[0] == 0 && mB[1] == 0)) {
            out[0] = out[1] = 0;
            return;
        }
        const int cnt = (rA == 0) + (rB == 0) + (rC == 0);
        if (cnt == 1) {
            const int *m = rA == 0 ? mA : (rB == 0 ? mB : mC);
            out[0] = m[0]; out[1] = m[1];
            return;
        }
        for (int c = 0; c < 2; c++) {
            const int a = mA[c], b = mB[c], cc = mC[c];
            out[c] = a + b + cc - std::min(a, std::min(b, cc)) - std::max(a, std::max(b, cc));
        }
    }

    void store_motion_mbaff(int addr, const h264g_mb &mb, bool intra) {
        for (int blk = 0; blk < 16; blk++) {
            Motion &m = mo_at(addr, blk);
            m = Motion();
            m.intra = intra ? 1 : 0;
            const int q = ((blk >> 3) << 1) | ((blk & 3) >> 1);
            for (int l = 0; l < 2; l++) {
                m.ref[l] = intra ? -1 : mb.ref_idx[l][q];
                m.mv[l][0] = intra ? 0 : mb.mv[l][blk][0];
                m.mv[l][1] = intra ? 0 : mb.mv[l][blk][1];
            }
        }
    }

    // Pairs in raster order, top MB then bottom MB.  Both classes are drawn first: a pair
    // of two P_Skip MBs carries no mb_field_decoding_flag, which is then inferred from the
    // left pair, else the pair above, else frame (7.4.4, H264SliceData.cpp:166-370).
    void run_mbaff() {
        motion.assign((size_t)W * H * 16, Motion());
        fld.assign((size_t)W * H, 0);
        for (int py = 0; py < H / 2; py++)
            for (int px = 0; px < W; px++) {
                const int p = py * W + px, a0 = 2 * p;
                int field = R.permille(500) ? 1 : 0;
                h264g_mb &t = pic.mbs[a0], &b = pic.mbs[a0 + 1];
                init_mb(t);
                init_mb(b);
                const int ct = choose_class(t), cb = choose_class(b);
                if (ct == H264G_MB_P_SKIP && cb == H264G_MB_P_SKIP)
                    field = px > 0 ? fld[2 * (p - 1)] : (py > 0 ? fld[2 * (p - W)] : 0);
                fld[a0] = fld[a0 + 1] = (uint8_t)field;
                for (int k = 0; k < 2; k++) {
                    const int addr = a0 + k, cls = k ? cb : ct;
                    h264g_mb &mb = k ? b : t;
                    if (field) mb.flags |= H264G_MBF_FIELD;
                    // intra neighbours (6.4.12.2): a frame bottom MB has its top MB above it and
                    // the left pair above-left; every other MB the pairs around its own
                    const bool A = px > 0, pb = py > 0;
                    const bool B = (!field && k) ? true : pb;
                    const bool Dn = (!field && k) ? A : (pb && px > 0);
                    gen_body(mb, cls, A, B, Dn, field != 0, [&](int mv[2]) { pskip_mv_mbaff(addr, mv); },
                             [](h264g_mb &) {});
                    store_motion_mbaff(addr, mb, cls <= H264G_MB_IPCM);
                }
            }
    }

    // Draws partitions, prediction lists, reference indices and final MVs.  The reference
    // indices of a field MB of an MBAFF frame range over the field list (2 x frames).
    void gen_inter(h264g_mb &mb, bool field) {
        const bool isB = mb.mb_type == H264G_MB_B;
        auto draw_pred = [&]() -> int {  // 1 = L0, 2 = L1, 3 = Bi
            if (!isB) return 1;
            int u = R.uni(0, 99);
            return u < 40 ? 3 : (u < 70 ? 1 : 2);
        };
        auto draw_ref = [&](int l) -> int { return R.uni(0, D.num_ref_idx_active[l] * (field ? 2 : 1) - 1); };
        // With implicit weights a bi-predicted partition must not use the same picture
        // from both lists: the reference divides by td = POC1 - POC0 = 0 there
        // (H264InterPrediction.cpp:2958) before its 32/32 fallback test.
        auto draw_pred_refs = [&](int &pred, int ref[2]) {
            pred = draw_pred();
            for (int l = 0; l < 2; l++) ref[l] = (pred & (1 << l)) ? draw_ref(l) : -1;
            if (pred == 3 && D.weighted_bipred_idc == 2) {
                for (int t = 0; t < 8 && D.ref_slot[0][ref[0]] == D.ref_slot[1][ref[1]]; t++) ref[1] = draw_ref(1);
                if (D.ref_slot[0][ref[0]] == D.ref_slot[1][ref[1]]) { pred = 1; ref[1] = -1; }
            }
        };
        // rectangles in 4x4 units: (x, y, w, h) with their quadrant list
        auto fill = [&](int x0, int y0, int w, int h, int l, int ref, int mvx, int mvy) {
            for (int y = y0; y < y0 + h; y++)
                for (int x = x0; x < x0 + w; x++) {
                    mb.mv[l][y * 4 + x][0] = (int16_t)mvx;
                    mb.mv[l][y * 4 + x][1] = (int16_t)mvy;
                }
            for (int y = y0; y < y0 + h; y += 2)
                for (int x = x0; x < x0 + w; x += 2) mb.ref_idx[l][(y / 2) * 2 + x / 2] = (int8_t)ref;
        };
        int shape = mb.part_shape;
        if (shape != H264G_PART_8x8) {
            int np = shape == H264G_PART_16x16 ? 1 : 2;
            for (int p = 0; p < np; p++) {
                int x0 = 0, y0 = 0, w = 4, h = 4;
                if (shape == H264G_PART_16x8) { h = 2; y0 = 2 * p; }
                if (shape == H264G_PART_8x16) { w = 2; x0 = 2 * p; }
                int pred, refs[2];
                draw_pred_refs(pred, refs);
                for (int l = 0; l < 2; l++)
                    if (pred & (1 << l)) {
                        int mvx = draw_mv_comp(), mvy = draw_mv_comp();
                        fill(x0, y0, w, h, l, refs[l], mvx, mvy);
                    }
            }
        } else {
            for (int q = 0; q < 4; q++) {
                if (mb.flags & H264G_MBF_DIRECT(q)) {  // direct: derived after the others' syntax
                    mb.sub_part_shape[q] = H264G_SUB_8x8;
                    continue;
                }
                int sub = R.uni(0, 3);
                mb.sub_part_shape[q] = (uint8_t)sub;
                int pred, refs[2];
                draw_pred_refs(pred, refs);
                int qx = (q % 2) * 2, qy = (q / 2) * 2;
                int sw = (sub == H264G_SUB_8x8 || sub == H264G_SUB_8x4) ? 2 : 1;
                int sh = (sub == H264G_SUB_8x8 || sub == H264G_SUB_4x8) ? 2 : 1;
                for (int l = 0; l < 2; l++) {
                    if (!(pred & (1 << l))) continue;
                    int ref = refs[l];
                    for (int sy = 0; sy < 2; sy += sh)
                        for (int sx = 0; sx < 2; sx += sw) {
                            int mvx = draw_mv_comp(), mvy = draw_mv_comp();
                            fill(qx + sx, qy + sy, sw, sh, l, ref, mvx, mvy);
                        }
                }
            }
        }
    }
};

uint64_t fnv1a(uint64_t h, const void *data, size_t n) {
    const uint8_t *p = (const uint8_t *)data;
    for (size_t i = 0; i < n; i++) { h ^= p[i]; h *= 0x100000001B3ull; }
    return h;
}

}  // namespace

extern "C" {

int32_t h264g_gen_default_params(int32_t config, h264g_gen_params *o) {
    if (!o) return H264G_E_ARG;
    std::memset(o, 0, sizeof(*o));
    o->config = config;
    o->seed = 0x5EED0000ull + (uint64_t)config;
    o->max_refs = 1;
    o->mv_mode = 0; o->mv_range = 64; o->mv_clamp_lo = -512; o->mv_clamp_hi = 511;
    o->qp_lo = 20; o->qp_hi = 36;
    o->zero_permille = 700; o->laplace_b_x100 = 200;
    o->skip_permille = 100;
    if (config == 1) {
        o->width_mbs = 22; o->height_mbs = 18; o->n_frames = 2;
        o->b_frames = 0; o->p_pattern = 1;
        return H264G_OK;
    }
    if (config == 2 || config == 0 || config == 3) {
        o->width_mbs = 120; o->height_mbs = 68; o->n_frames = 8;
        o->b_frames = 2; o->max_refs = 4; o->t8x8 = 1;
        o->mv_mode = 1; o->mv_range = 24;
        o->qp_lo = 18; o->qp_hi = 40;
        o->zero_permille = 600; o->laplace_b_x100 = 300;
        o->intra_in_inter_permille = 50;
        o->weighted_pred_flag = 0; o->weighted_bipred_idc = 1;
        o->filter_offsets = 1;
        if (config == 3) {  // MBAFF frames, 4x4 transform only
            o->t8x8 = 0;
            o->mbaff = 1;
        }
        return H264G_OK;
    }
    return H264G_E_ARG;
}

int32_t h264g_gen_create(const h264g_gen_params *params, h264g_gen **out) {
    if (!params || !out) return H264G_E_ARG;
    const h264g_gen_params &p = *params;
    if (p.width_mbs <= 0 || p.height_mbs <= 0 || p.n_frames <= 0 || p.max_refs < 1 || p.max_refs > 16 ||
        p.qp_lo < 0 || p.qp_hi > 51 || p.qp_lo > p.qp_hi || p.qp_hi - p.qp_lo > 25 || p.spare_slots < 0 ||
        p.spare_slots > 16)
        return h264g_set_error(H264G_E_ARG, "gen: bad params");
    // MBAFF: pairs of MB rows; implicit weights of field MBs would need field POCs
    if (p.mbaff && (p.height_mbs % 2 || p.weighted_bipred_idc == 2 || p.constrained_intra_pred || p.direct_permille))
        return h264g_set_error(H264G_E_ARG,
                               "gen: MBAFF needs an even height, no implicit weights, no constrained intra, no direct");
    // direct MBs take both lists' indices from their neighbours: with implicit weights one
    // picture in both lists would make the reference divide by td = 0
    // (H264InterPrediction.cpp:2958)
    if (p.direct_permille < 0 || p.direct_permille > 1000 || (p.direct_permille && p.weighted_bipred_idc == 2))
        return h264g_set_error(H264G_E_ARG, "gen: direct MBs need direct_permille in 0..1000 and no implicit weights");
    if (p.slices < 0 || p.slices > p.width_mbs * p.height_mbs || (p.slices > 1 && p.mbaff))
        return h264g_set_error(H264G_E_ARG, "gen: several slices need frame MBs (no MBAFF)");
    h264g_gen *g = new h264g_gen();
    g->p = p;
    std::vector<PicPlan> plan = plan_gop(p.n_frames, p.b_frames, p.gop);

    // DPB bookkeeping: reference frames in decoding order with their slot, POC, frame_num.
    struct RefEntry { int slot, poc, frame_num, pic; };
    std::vector<RefEntry> refs;
    int n_slots = p.max_refs + 1 + p.spare_slots, last_dst = n_slots - 1;
    std::vector<int> slot_busy(n_slots, 0);
    g->dpb_slots = n_slots;
    g->pics.resize(plan.size());
    for (size_t i = 0; i < plan.size(); i++) {
        const PicPlan &pp = plan[i];
        if (pp.idr) { refs.clear(); std::fill(slot_busy.begin(), slot_busy.end(), 0); }
        Picture &pic = g->pics[i];
        h264g_pic_desc &d = pic.desc;
        std::memset(&d, 0, sizeof(d));
        Rng rng(p.seed * 0x100000001B3ull + (uint64_t)i * 0x9E3779B97F4A7C15ull + 1);
        d.slice_type = pp.slice_type;
        d.structure = H264G_FRAME;
        d.mbaff = p.mbaff ? 1 : 0;
        d.constrained_intra_pred_flag = p.constrained_intra_pred ? 1 : 0;
        d.poc = 2 * pp.display;
        d.disable_deblocking_filter_idc = p.disable_deblocking_filter_idc;
        if (p.filter_offsets) {
            d.filter_offset_a = 2 * rng.uni(-2, 2);
            d.filter_offset_b = 2 * rng.uni(-2, 2);
        }
        d.chroma_qp_index_offset = p.chroma_qp_index_offset;
        d.second_chroma_qp_index_offset = p.second_chroma_qp_index_offset;
        d.weighted_pred_flag = p.weighted_pred_flag;
        d.weighted_bipred_idc = p.weighted_bipred_idc;
        for (int l = 0; l < 2; l++)
            for (int r = 0; r < 32; r++) d.ref_slot[l][r] = -1;

        // ---- reference lists (8.2.4.2.1 / 8.2.4.2.3 for frames)
        std::vector<RefEntry> l0, l1;
        if (pp.slice_type == H264G_SLICE_P) {
            l0 = refs;
            std::sort(l0.begin(), l0.end(), [](const RefEntry &a, const RefEntry &b) { return a.frame_num > b.frame_num; });
        } else if (pp.slice_type == H264G_SLICE_B) {
            std::vector<RefEntry> before, after;
            for (auto &r : refs) (r.poc < d.poc ? before : after).push_back(r);
            std::sort(before.begin(), before.end(), [](const RefEntry &a, const RefEntry &b) { return a.poc > b.poc; });
            std::sort(after.begin(), after.end(), [](const RefEntry &a, const RefEntry &b) { return a.poc < b.poc; });
            l0 = before; l0.insert(l0.end(), after.begin(), after.end());
            l1 = after; l1.insert(l1.end(), before.begin(), before.end());
            if (l1.size() > 1 && l0.size() == l1.size()) {
                bool same = true;
                for (size_t k = 0; k < l0.size(); k++) same &= l0[k].slot == l1[k].slot;
                if (same) std::swap(l1[0], l1[1]);
            }
        }
        std::vector<RefEntry> *lists[2] = {&l0, &l1};
        for (int l = 0; l < 2; l++) {
            int n = (int)std::min<size_t>(lists[l]->size(), (size_t)p.max_refs);
            d.num_ref_idx_active[l] = n;
            for (int r = 0; r < n; r++) {
                d.ref_slot[l][r] = (*lists[l])[r].slot;
                d.ref_poc[l][r] = (*lists[l])[r].poc;
            }
        }
        // ---- explicit weighted prediction tables (pred_weight_table syntax)
        bool explicit_wp = (pp.slice_type == H264G_SLICE_P && p.weighted_pred_flag == 1) ||
                           (pp.slice_type == H264G_SLICE_B && p.weighted_bipred_idc == 1);
        if (explicit_wp) {
            d.luma_log2_weight_denom = 5;
            d.chroma_log2_weight_denom = 5;
            int nl = pp.slice_type == H264G_SLICE_B ? 2 : 1;
            for (int l = 0; l < nl; l++) {
                for (int r = 0; r < d.num_ref_idx_active[l]; r++) {
                    d.luma_weight[l][r] = (int16_t)rng.uni(16, 48);
                    d.luma_offset[l][r] = (int16_t)rng.uni(-8, 8);
                    for (int c = 0; c < 2; c++) {
                        d.chroma_weight[l][r][c] = (int16_t)rng.uni(16, 48);
                        d.chroma_offset[l][r][c] = (int16_t)rng.uni(-8, 8);
                    }
                }
                d.luma_weight_flag_last[l] = d.num_ref_idx_active[l] > 0 ? 1 : 0;
            }
        }
        // ---- destination slot
        int dst = -1;
        for (int k = 0; k < n_slots; k++) {  // lowest free slot, or (spare slots) the next free one
            const int s = p.spare_slots > 0 ? (last_dst + 1 + k) % n_slots : k;
            if (!slot_busy[s]) { dst = s; break; }
        }
        if (dst < 0) { delete g; return h264g_set_error(H264G_E_STATE, "gen: no free DPB slot"); }
        d.dst_slot = dst;
        last_dst = dst;

        // ---- macroblocks
        const h264g_mb *col = pp.slice_type == H264G_SLICE_B && !l1.empty() ? g->pics[l1[0].pic].mbs.data() : nullptr;
        if (p.direct_permille && pp.slice_type == H264G_SLICE_B && !col) {
            delete g;
            return h264g_set_error(H264G_E_STATE, "gen: B picture without RefPicList1[0]");
        }
        PictureGen pg(p, rng, d, pic, col);
        pg.run();

        // ---- reference marking: sliding window over max_refs frames
        if (pp.is_ref) {
            if ((int)refs.size() == p.max_refs) {
                auto oldest = std::min_element(refs.begin(), refs.end(), [](const RefEntry &a, const RefEntry &b) { return a.frame_num < b.frame_num; });
                slot_busy[oldest->slot] = 0;
                refs.erase(oldest);
            }
            refs.push_back({dst, d.poc, pp.frame_num, (int)i});
            slot_busy[dst] = 1;
        }
    }
    uint64_t h = 0xCBF29CE484222325ull;
    for (auto &pic : g->pics) {
        // the descriptor up to the scaling matrices, which the generator leaves flat (the
        // digest of a stream stays what it was before the descriptor carried them)
        h = fnv1a(h, &pic.desc, offsetof(h264g_pic_desc, scaling_matrix));
        h = fnv1a(h, pic.mbs.data(), pic.mbs.size() * sizeof(h264g_mb));
        h = fnv1a(h, pic.levels.data(), pic.levels.size() * sizeof(int16_t));
    }
    g->digest = h;
    *out = g;
    return H264G_OK;
}

int32_t h264g_gen_destroy(h264g_gen *gen) {
    delete gen;
    return H264G_OK;
}

int32_t h264g_gen_num_pictures(const h264g_gen *gen) { return gen ? (int32_t)gen->pics.size() : H264G_E_ARG; }

int32_t h264g_gen_dpb_slots(const h264g_gen *gen) { return gen ? gen->dpb_slots : H264G_E_ARG; }

uint64_t h264g_gen_digest(const h264g_gen *gen) { return gen ? gen->digest : 0; }

int32_t h264g_gen_picture(h264g_gen *gen, int32_t idx, h264g_pic_desc *desc, const h264g_mb **mbs,
                          uint32_t *n_mbs, const int16_t **levels, uint64_t *n_levels) {
    if (!gen || idx < 0 || idx >= (int32_t)gen->pics.size()) return H264G_E_ARG;
    const Picture &p = gen->pics[idx];
    if (desc) *desc = p.desc;
    if (mbs) *mbs = p.mbs.data();
    if (n_mbs) *n_mbs = (uint32_t)p.mbs.size();
    if (levels) *levels = p.levels.data();
    if (n_levels) *n_levels = p.levels.size();
    return H264G_OK;
}

}  // extern "C"
