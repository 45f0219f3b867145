// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY (never shipped, never on the product path).
//
// Drives the reference decoder's own reconstruction code (CH264PictureBase methods
// of jfu222/h264_video_decoder_demo, compiled from /root/reference by
// oracle/Makefile into oracle/_ref/) with synthetic syntax from the h264g
// generator, and writes the reconstructed pictures.  Those pictures are the golden
// vectors the oracle restatement (oracle/refcpu.c) and the HIP engine are pinned to.
//
// Per macroblock the harness fills CH264MacroBlock exactly as macroblock_layer()
// and residual() would after entropy decoding (H264MacroBlock.cpp:912-1139,
// 1592-1856), then calls the same reconstruction dispatch as
// CH264SliceData::slice_data (H264SliceData.cpp:386-487; P_Skip: 222-227), and
// ends each picture with Deblocking_filter_process (H264PictureBase.cpp:707).
//
// Syntax elements that the reference predicts (mvd, prev_intra*_pred_mode_flag,
// rem_intra*_pred_mode) are solved with the reference's own predictors so that the
// reconstructed final values equal the generator's:
//   * MVs: each partition is decoded with mvd = 0, the resulting MV is the
//     reference's mvp, and mvd = target - mvp; partitions are solved in decoding
//     order with the MB's motion state cleared before each run (H264InterPrediction.cpp:412-667).
//   * intra modes: getIntra4x4PredMode / getIntra8x8PredMode with the flag set give
//     predIntraNxNPredMode; the flag / rem pair then encodes the target mode.
// P_Skip MVs are not solvable (no mvd): the reference derives them and the harness
// checks them against the generator's record.
//
// Usage: ref_harness --config N [--frames F --width W --height H --seed S]
//                    [--ipcm P] [--wp P,B] [--cqp a,b] [--deblock idc]
//                    --out DIR [--predeblock] [--bgr] [--mvd] [--time]
// --bgr also writes every picture converted by the reference's
// CH264PictureBase::convertYuv420pToBgr24 (rows padded to 4 bytes) as bgr_NNN.raw.
// --mvd also writes the solved mvd of every picture as mvd_NNN.bin: int16
// [MB][list][16][2], each (sub-)partition's mvd at its top-left 4x4 block (raster index
// inside the MB, the layout of h264g_mb::mv) for the lists it uses, zero elsewhere.
#include "H264PicturesGOP.h"
#include "H264Picture.h"
#include "H264PictureBase.h"
#include "H264MacroBlock.h"
#include "H264SliceHeader.h"
#include "h264g.h"

#include <malloc.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

extern MB_TYPE_I_SLICES_T mb_type_I_slices_define[27];
extern MB_TYPE_P_SP_SLICES_T mb_type_P_SP_slices_define[6];
extern MB_TYPE_B_SLICES_T mb_type_B_slices_define[24];
extern SUB_MB_TYPE_P_MBS_T sub_mb_type_P_mbs_define[4];
extern SUB_MB_TYPE_B_MBS_T sub_mb_type_B_mbs_define[13];



namespace {

int fail(const char *msg) {
    fprintf(stderr, "ref_harness: %s\n", msg);
    exit(2);
}

// ----------------------------------------------------------- levels
// Expands one MB's mask/value level stream into dense scan-order arrays.
struct DenseLevels {
    int32_t dc16[16];
    int32_t ac16[16][16];   // Intra16x16ACLevel: index k-1 holds scan position k
    int32_t l4[16][16];     // LumaLevel4x4
    int32_t l8[4][64];      // LumaLevel8x8
    int32_t cdc[2][16];
    int32_t cac[2][16][16]; // ChromaACLevel: index k-1 holds scan position k
    int32_t pcm[384];
};

void expand_levels(const h264g_mb &mb, const int16_t *lv, DenseLevels &d) {
    memset(&d, 0, sizeof(d));
    const int16_t *p = lv + mb.level_offset;
    if (mb.mb_type == H264G_MB_IPCM) {
        for (int i = 0; i < 384; i++) d.pcm[i] = p[i];
        return;
    }
    int cbp_l = mb.cbp & 15, cbp_c = mb.cbp >> 4;
    bool i16 = mb.mb_type == H264G_MB_I16x16, t8 = (mb.flags & H264G_MBF_T8x8) != 0;
    // collect (dst pointer, n, first) per mask in stream order
    struct Blk { int32_t *dst; int n; int shift; };  // dst[k - shift] = value at scan position k
    std::vector<Blk> blks;
    if (i16) blks.push_back({d.dc16, 16, 0});
    for (int q = 0; q < 4; q++) {
        if (!(cbp_l & (1 << q))) continue;
        if (t8) for (int w = 0; w < 4; w++) blks.push_back({&d.l8[q][w * 16], 16, 0});
        else
            for (int b = 0; b < 4; b++) {
                if (i16) blks.push_back({d.ac16[q * 4 + b], 16, 1});
                else blks.push_back({d.l4[q * 4 + b], 16, 0});
            }
    }
    if (cbp_c >= 1) for (int c = 0; c < 2; c++) blks.push_back({d.cdc[c], 4, 0});
    if (cbp_c == 2)
        for (int c = 0; c < 2; c++)
            for (int b = 0; b < 4; b++) blks.push_back({d.cac[c][b], 16, 1});
    const int16_t *vals = p + blks.size();
    for (auto &b : blks) {
        uint16_t m = (uint16_t)p[&b - &blks[0]];
        for (int k = 0; k < b.n; k++)
            if (m & (1u << k)) {
                if (k < b.shift) fail("level at a position the block does not carry");
                b.dst[k - b.shift] = *vals++;
            }
    }
}

int total_coeff(const int32_t *v, int n) {
    int c = 0;
    for (int i = 0; i < n; i++) c += v[i] != 0;
    return c;
}

// --------------------------------------------------------- mb types
H264_MB_TYPE b_16x16_name(int pred) {
    return pred == 1 ? B_L0_16x16 : (pred == 2 ? B_L1_16x16 : B_Bi_16x16);
}
int b_rect_mb_type(int shape, int p0, int p1) {  // Table 7-14 mb_type for 16x8 / 8x16
    // pred codes: 1 = L0, 2 = L1, 3 = Bi
    static const int t16x8[4][4] = {{0, 0, 0, 0}, {0, 4, 8, 12}, {0, 10, 6, 14}, {0, 16, 18, 20}};
    int v = t16x8[p0][p1];
    return shape == H264G_PART_16x8 ? v : v + 1;
}
int b_sub_type(int sub, int pred) {  // Table 7-18 sub_mb_type
    static const int base[4][4] = {{0, 1, 2, 3}, {0, 4, 6, 8}, {0, 5, 7, 9}, {0, 10, 11, 12}};
    return base[sub][pred];
}
int quad_pred(const h264g_mb &mb, int q) {
    return (mb.ref_idx[0][q] >= 0 ? 1 : 0) | (mb.ref_idx[1][q] >= 0 ? 2 : 0);
}

struct Harness {
    h264g_gen *gen = nullptr;
    h264g_gen_params gp;
    int W = 0, H = 0;
    std::vector<CH264Picture *> slots;
    CH264Picture *dpb[16];
    CH264SliceHeader sh;
    bool predeblock = false;
    bool mvd = false;
    bool bgr = false;
    std::string out;
    double recon_seconds = 0;
    long mismatches = 0;

    void setup_headers() {
        CH264SPS &sps = sh.m_sps;
        sps.chroma_format_idc = 1;
        sps.separate_colour_plane_flag = 0;
        sps.Chroma_Format = 1;  // CHROMA_FORMAT_IDC_420
        sps.ChromaArrayType = 1;
        sps.SubWidthC = 2; sps.SubHeightC = 2;
        sps.MbWidthC = 8; sps.MbHeightC = 8;
        sps.BitDepthY = 8; sps.BitDepthC = 8;
        sps.QpBdOffsetY = 0; sps.QpBdOffsetC = 0;
        sps.PicWidthInMbs = W;
        sps.FrameHeightInMbs = H;
        sps.PicHeightInMapUnits = H;
        sps.PicSizeInMapUnits = W * H;
        sps.PicWidthInSamplesL = W * 16;
        sps.PicWidthInSamplesC = W * 8;
        sps.frame_mbs_only_flag = gp.mbaff ? 0 : 1;
        sps.mb_adaptive_frame_field_flag = gp.mbaff ? 1 : 0;
        sps.direct_8x8_inference_flag = 1;
        sps.qpprime_y_zero_transform_bypass_flag = 0;
        sps.max_num_ref_frames = gp.max_refs;
        CH264PPS &pps = sh.m_pps;
        pps.entropy_coding_mode_flag = 0;
        pps.transform_8x8_mode_flag = gp.t8x8;
        pps.constrained_intra_pred_flag = 0;
        pps.deblocking_filter_control_present_flag = 1;
        for (int i = 0; i < 6; i++) {
            for (int k = 0; k < 16; k++) sh.ScalingList4x4[i][k] = 16;  // Flat_4x4_16
            for (int k = 0; k < 64; k++) sh.ScalingList8x8[i][k] = 16;  // Flat_8x8_16
        }
        sh.field_pic_flag = 0;
        sh.bottom_field_flag = 0;
        sh.MbaffFrameFlag = gp.mbaff ? 1 : 0;
        sh.PicHeightInMbs = H;
        sh.PicHeightInSamplesL = H * 16;
        sh.PicHeightInSamplesC = H * 8;
        sh.PicSizeInMbs = W * H;
        sh.m_picture_coded_type = H264_PICTURE_CODED_TYPE_FRAME;
    }

    void set_slice(const h264g_pic_desc &d) {
        sh.slice_type = d.slice_type;
        sh.slice_type_fixed = d.slice_type;
        sh.disable_deblocking_filter_idc = d.disable_deblocking_filter_idc;
        sh.slice_alpha_c0_offset_div2 = d.filter_offset_a / 2;
        sh.slice_beta_offset_div2 = d.filter_offset_b / 2;
        sh.FilterOffsetA = d.filter_offset_a;
        sh.FilterOffsetB = d.filter_offset_b;
        sh.m_pps.chroma_qp_index_offset = d.chroma_qp_index_offset;
        sh.m_pps.second_chroma_qp_index_offset = d.second_chroma_qp_index_offset;
        sh.m_pps.constrained_intra_pred_flag = d.constrained_intra_pred_flag;
        sh.m_pps.weighted_pred_flag = d.weighted_pred_flag;
        sh.m_pps.weighted_bipred_idc = d.weighted_bipred_idc;
        sh.num_ref_idx_l0_active_minus1 = d.num_ref_idx_active[0] > 0 ? d.num_ref_idx_active[0] - 1 : 0;
        sh.num_ref_idx_l1_active_minus1 = d.num_ref_idx_active[1] > 0 ? d.num_ref_idx_active[1] - 1 : 0;
        sh.direct_spatial_mv_pred_flag = 1;  // spatial direct (B_Skip / B_Direct_* records)
        // pred_weight_table as left by parsing (H264SliceHeader.cpp:586-665)
        sh.luma_log2_weight_denom = d.luma_log2_weight_denom;
        sh.chroma_log2_weight_denom = d.chroma_log2_weight_denom;
        sh.luma_weight_l0_flag = d.luma_weight_flag_last[0];
        sh.luma_weight_l1_flag = d.luma_weight_flag_last[1];
        for (int i = 0; i < 32; i++) {
            sh.luma_weight_l0[i] = d.luma_weight[0][i];
            sh.luma_offset_l0[i] = d.luma_offset[0][i];
            sh.luma_weight_l1[i] = d.luma_weight[1][i];
            sh.luma_offset_l1[i] = d.luma_offset[1][i];
            for (int c = 0; c < 2; c++) {
                sh.chroma_weight_l0[i][c] = d.chroma_weight[0][i][c];
                sh.chroma_offset_l0[i][c] = d.chroma_offset[0][i][c];
                sh.chroma_weight_l1[i][c] = d.chroma_weight[1][i][c];
                sh.chroma_offset_l1[i][c] = d.chroma_offset[1][i][c];
            }
        }
    }

    // CurrMbAddr -> (mb_x, mb_y) as CH264SliceData::slice_data sets them (H264SliceData.cpp:
    // 176-179): in MBAFF frames mb_y counts MB rows of pairs, 2 * pair row + bottom.
    void set_position(CH264PictureBase &pic, int addr, const h264g_mb &r) {
        const int mbaff = gp.mbaff ? 1 : 0, w2 = W * (1 + mbaff);
        pic.CurrMbAddr = addr;
        pic.mb_x = (addr % w2) / (1 + mbaff);
        pic.mb_y = (addr / w2) * (1 + mbaff) + (addr % w2) % (1 + mbaff);
        // the slice-level copy of mb_field_decoding_flag, read by the interpolation's
        // effective reference height (H264InterPrediction.cpp:2351-2358, 2492-2499)
        pic.m_h264_slice_data.mb_field_decoding_flag = (r.flags & H264G_MBF_FIELD) ? 1 : 0;
    }

    // Fills the MB fields macroblock_layer() would set (without entropy decoding).
    void fill_mb(CH264PictureBase &pic, int addr, const h264g_mb &r, const DenseLevels &lv, int qp_prev) {
        CH264MacroBlock &m = pic.m_mbs[addr];
        m.field_pic_flag = 0;
        m.bottom_field_flag = 0;
        m.mb_skip_flag = 0;
        m.mb_field_decoding_flag = (r.flags & H264G_MBF_FIELD) ? 1 : 0;
        m.MbaffFrameFlag = gp.mbaff ? 1 : 0;
        m.constrained_intra_pred_flag = sh.m_pps.constrained_intra_pred_flag;
        m.disable_deblocking_filter_idc = sh.disable_deblocking_filter_idc;
        m.CurrMbAddr = addr;
        m.slice_id = 0;
        m.slice_number = 1;
        m.m_slice_type = sh.slice_type;
        m.FilterOffsetA = sh.FilterOffsetA;
        m.FilterOffsetB = sh.FilterOffsetB;
        pic.Inverse_macroblock_scanning_process(m.MbaffFrameFlag, addr, m.mb_field_decoding_flag, m.m_mb_position_x,
                                                m.m_mb_position_y);

        const int st = sh.slice_type;
        int t8 = (r.flags & H264G_MBF_T8x8) ? 1 : 0;
        int cbp_l = r.cbp & 15, cbp_c = r.cbp >> 4;
        int intra = r.mb_type <= H264G_MB_IPCM;
        // raw mb_type in the slice's table, then fix_mb_type (H264MacroBlock.cpp:403)
        int raw = 0;
        if (intra) {
            int itype;
            if (r.mb_type == H264G_MB_I4x4 || r.mb_type == H264G_MB_I8x8) itype = 0;
            else if (r.mb_type == H264G_MB_IPCM) itype = 25;
            else itype = 1 + r.intra16x16_pred_mode + 4 * cbp_c + (cbp_l ? 12 : 0);
            raw = itype + (st == H264G_SLICE_P ? 5 : (st == H264G_SLICE_B ? 23 : 0));
        } else if (r.mb_type == H264G_MB_P_SKIP) {
            raw = 5;
        } else if (r.mb_type == H264G_MB_P) {
            raw = r.part_shape;  // 0 16x16, 1 16x8, 2 8x16, 3 P_8x8
        } else if ((r.flags & H264G_MBF_DIRECT_ALL) == H264G_MBF_DIRECT_ALL) {
            raw = (r.flags & H264G_MBF_SKIP) ? 23 : 0;  // B_Skip (inferred, as H264MacroBlock.cpp:1176) / B_Direct_16x16
        } else {
            if (r.part_shape == H264G_PART_16x16) raw = quad_pred(r, 0);                 // 1..3
            else if (r.part_shape == H264G_PART_8x8) raw = 22;
            else if (r.part_shape == H264G_PART_16x8) raw = b_rect_mb_type(r.part_shape, quad_pred(r, 0), quad_pred(r, 2));
            else raw = b_rect_mb_type(r.part_shape, quad_pred(r, 0), quad_pred(r, 1));
        }
        m.mb_type = raw;
        m.transform_size_8x8_flag = t8;
        if (r.mb_type == H264G_MB_P_SKIP || (r.flags & H264G_MBF_SKIP)) {
            // macroblock_layer_mb_skip (H264MacroBlock.cpp:1142-1204)
            m.m_slice_type_fixed = st;
            m.m_mb_type_fixed = raw;
            m.mb_qp_delta = 0;
        } else {
            CH264MacroBlock::fix_mb_type(st, raw, m.m_slice_type_fixed, m.m_mb_type_fixed);
        }
        CH264MacroBlock::MbPartPredMode(m.m_slice_type_fixed, t8, m.m_mb_type_fixed, 0, m.m_NumMbPart,
                                        m.CodedBlockPatternChroma, m.CodedBlockPatternLuma, m.Intra16x16PredMode,
                                        m.m_name_of_mb_type, m.m_mb_pred_mode);
        // set_mb_type_X_slice_info (H264MacroBlock.cpp:812)
        int sft = m.m_slice_type_fixed % 5;
        if (sft == H264_SLIECE_TYPE_I) {
            m.mb_type_I_slice = mb_type_I_slices_define[m.m_mb_type_fixed == 0 ? t8 : m.m_mb_type_fixed + 1];
        } else if (sft == H264_SLIECE_TYPE_P) {
            m.mb_type_P_SP_slice = mb_type_P_SP_slices_define[m.m_mb_type_fixed];
            m.MbPartWidth = m.mb_type_P_SP_slice.MbPartWidth;
            m.MbPartHeight = m.mb_type_P_SP_slice.MbPartHeight;
            m.m_NumMbPart = m.mb_type_P_SP_slice.NumMbPart;
        } else if (sft == H264_SLIECE_TYPE_B) {
            m.mb_type_B_slice = mb_type_B_slices_define[m.m_mb_type_fixed];
            m.MbPartWidth = m.mb_type_B_slice.MbPartWidth;
            m.MbPartHeight = m.mb_type_B_slice.MbPartHeight;
            m.m_NumMbPart = m.mb_type_B_slice.NumMbPart;
        }
        if (!intra && r.mb_type != H264G_MB_P_SKIP) {
            m.coded_block_pattern = r.cbp & 0x3F;
            m.CodedBlockPatternLuma = cbp_l;
            m.CodedBlockPatternChroma = cbp_c;
        } else if (r.mb_type == H264G_MB_I4x4 || r.mb_type == H264G_MB_I8x8) {
            m.coded_block_pattern = r.cbp & 0x3F;
            m.CodedBlockPatternLuma = cbp_l;
            m.CodedBlockPatternChroma = cbp_c;
        }
        // sub-macroblock types (sub_mb_pred, H264MacroBlock.cpp:1425-1473); none for B_Skip /
        // B_Direct_16x16, B_Direct_8x8 (0) for the direct quadrants of B_8x8
        const bool all_direct = (r.flags & H264G_MBF_DIRECT_ALL) == H264G_MBF_DIRECT_ALL;
        if (r.part_shape == H264G_PART_8x8 && (r.mb_type == H264G_MB_P || r.mb_type == H264G_MB_B) && !all_direct) {
            for (int q = 0; q < 4; q++) {
                int sub = r.sub_part_shape[q];
                if (r.mb_type == H264G_MB_P) {
                    m.sub_mb_type[q] = sub;
                    const SUB_MB_TYPE_P_MBS_T &t = sub_mb_type_P_mbs_define[sub];
                    m.m_name_of_sub_mb_type[q] = t.name_of_sub_mb_type;
                    m.m_sub_mb_pred_mode[q] = t.SubMbPredMode;
                    m.NumSubMbPart[q] = t.NumSubMbPart;
                    m.SubMbPartWidth[q] = t.SubMbPartWidth;
                    m.SubMbPartHeight[q] = t.SubMbPartHeight;
                } else {
                    int s = (r.flags & H264G_MBF_DIRECT(q)) ? 0 : b_sub_type(sub, quad_pred(r, q));
                    m.sub_mb_type[q] = s;
                    const SUB_MB_TYPE_B_MBS_T &t = sub_mb_type_B_mbs_define[s];
                    m.m_name_of_sub_mb_type[q] = t.name_of_sub_mb_type;
                    m.m_sub_mb_pred_mode[q] = t.SubMbPredMode;
                    m.NumSubMbPart[q] = t.NumSubMbPart;
                    m.SubMbPartWidth[q] = t.SubMbPartWidth;
                    m.SubMbPartHeight[q] = t.SubMbPartHeight;
                }
            }
        }
        // reference indices per partition
        if (!intra && r.mb_type != H264G_MB_P_SKIP) {
            for (int p = 0; p < 4; p++) {
                int q = p;
                if (r.part_shape == H264G_PART_16x8) q = p * 2;
                m.ref_idx_l0[p] = r.ref_idx[0][q < 4 ? q : 0] >= 0 ? r.ref_idx[0][q < 4 ? q : 0] : 0;
                m.ref_idx_l1[p] = r.ref_idx[1][q < 4 ? q : 0] >= 0 ? r.ref_idx[1][q < 4 ? q : 0] : 0;
            }
        }
        m.intra_chroma_pred_mode = intra ? r.intra_chroma_pred_mode : 0;
        // QPY (7-37)
        m.mb_qp_delta = (int32_t)r.qp_y - qp_prev;
        m.QPY = r.qp_y;
        m.QP1Y = r.qp_y;
        m.TransformBypassModeFlag = 0;
        // residual (H264MacroBlock.cpp:1592-1856) with CAVLC bookkeeping
        memcpy(m.Intra16x16DCLevel, lv.dc16, sizeof(lv.dc16));
        memcpy(m.Intra16x16ACLevel, lv.ac16, sizeof(lv.ac16));
        memcpy(m.LumaLevel4x4, lv.l4, sizeof(lv.l4));
        memcpy(m.LumaLevel8x8, lv.l8, sizeof(lv.l8));
        memcpy(m.ChromaDCLevel, lv.cdc, sizeof(lv.cdc));
        memcpy(m.ChromaACLevel, lv.cac, sizeof(lv.cac));
        if (r.mb_type == H264G_MB_IPCM) {
            for (int i = 0; i < 256; i++) m.pcm_sample_luma[i] = lv.pcm[i];
            for (int i = 0; i < 128; i++) m.pcm_sample_chroma[i] = lv.pcm[256 + i];
        }
        bool has_residual = cbp_l || cbp_c || r.mb_type == H264G_MB_I16x16;
        if (has_residual) {
            bool i16 = r.mb_type == H264G_MB_I16x16;
            if (i16) m.mb_luma_4x4_non_zero_count_coeff[0] = (uint8_t)total_coeff(lv.dc16, 16);
            for (int q = 0; q < 4; q++)
                for (int b = 0; b < 4; b++) {
                    int blk = q * 4 + b;
                    int32_t tmp4[16];
                    const int32_t *src;
                    int n = 16;
                    if (t8) {  // CAVLC 8x8: the four interleaved 4x4 blocks (H264MacroBlock.cpp:1818-1825)
                        for (int i = 0; i < 16; i++) tmp4[i] = lv.l8[q][4 * i + b];
                        src = tmp4;
                        m.LumaLevel4x4[blk][0] = 0;
                        for (int i = 0; i < 16; i++) m.LumaLevel4x4[blk][i] = tmp4[i];
                    } else if (i16) { src = lv.ac16[blk]; n = 15; }
                    else src = lv.l4[blk];
                    if (cbp_l & (1 << q)) {
                        m.mb_luma_4x4_non_zero_count_coeff[blk] = (uint8_t)total_coeff(src, n);
                        m.mb_luma_8x8_non_zero_count_coeff[q] += m.mb_luma_4x4_non_zero_count_coeff[blk];
                    }
                    if (t8) m.mb_luma_8x8_non_zero_count_coeff[q] += m.mb_luma_4x4_non_zero_count_coeff[blk];
                }
            for (int c = 0; c < 2; c++) {
                if (cbp_c & 3) m.mb_chroma_4x4_non_zero_count_coeff[c][0] = (uint8_t)total_coeff(lv.cdc[c], 4);
                if (cbp_c & 2)
                    for (int b = 0; b < 4; b++) m.mb_chroma_4x4_non_zero_count_coeff[c][b] = (uint8_t)total_coeff(lv.cac[c][b], 15);
            }
        }
    }

    void clear_motion(CH264MacroBlock &m) {
        memset(m.m_MvL0, 0, sizeof(m.m_MvL0));
        memset(m.m_MvL1, 0, sizeof(m.m_MvL1));
        memset(m.m_RefIdxL0, 0, sizeof(m.m_RefIdxL0));
        memset(m.m_RefIdxL1, 0, sizeof(m.m_RefIdxL1));
        memset(m.m_PredFlagL0, 0, sizeof(m.m_PredFlagL0));
        memset(m.m_PredFlagL1, 0, sizeof(m.m_PredFlagL1));
        memset(m.m_isDecoded, 0, sizeof(m.m_isDecoded));
    }

    // (mbPartIdx, subMbPartIdx) -> top-left 4x4 block of the (sub-)partition, raster index
    static int part_block(const h264g_mb &r, int part, int sub) {
        int x = 0, y = 0;
        switch (r.part_shape) {
        case H264G_PART_16x16: break;
        case H264G_PART_16x8: y = part * 2; break;
        case H264G_PART_8x16: x = part * 2; break;
        default: {
            x = (part % 2) * 2; y = (part / 2) * 2;
            int s = r.sub_part_shape[part];
            if (s == H264G_SUB_8x4) y += sub;
            else if (s == H264G_SUB_4x8) x += sub;
            else if (s == H264G_SUB_4x4) { x += sub % 2; y += sub / 2; }
        }
        }
        return y * 4 + x;
    }

    int decode_mb(CH264PictureBase &pic, int addr, const h264g_mb &r) {
        CH264MacroBlock &m = pic.m_mbs[addr];
        int ret = 0;
        uint8_t *Y = pic.m_pic_buff_luma, *Cb = pic.m_pic_buff_cb, *Cr = pic.m_pic_buff_cr;
        int wl = pic.PicWidthInSamplesL, wc = pic.PicWidthInSamplesC;
        if (r.mb_type == H264G_MB_P_SKIP || (r.flags & H264G_MBF_SKIP)) return pic.Inter_prediction_process();
        switch (m.m_mb_pred_mode) {
        case Intra_4x4:
            ret |= pic.transform_decoding_process_for_4x4_luma_residual_blocks(0, 0, 8, wl, Y);
            ret |= pic.transform_decoding_process_for_chroma_samples(1, wc, Cb);
            ret |= pic.transform_decoding_process_for_chroma_samples(0, wc, Cr);
            return ret;
        case Intra_8x8:
            ret |= pic.transform_decoding_process_for_8x8_luma_residual_blocks(0, 0, 8, wl, m.LumaLevel8x8, Y);
            ret |= pic.transform_decoding_process_for_chroma_samples(1, wc, Cb);
            ret |= pic.transform_decoding_process_for_chroma_samples(0, wc, Cr);
            return ret;
        case Intra_16x16:
            ret |= pic.transform_decoding_process_for_luma_samples_of_Intra_16x16_macroblock_prediction_mode(
                0, 8, m.QP1Y, wl, m.Intra16x16DCLevel, m.Intra16x16ACLevel, Y);
            ret |= pic.transform_decoding_process_for_chroma_samples(1, wc, Cb);
            ret |= pic.transform_decoding_process_for_chroma_samples(0, wc, Cr);
            return ret;
        default:
            break;
        }
        if (m.m_name_of_mb_type == I_PCM) return pic.Sample_construction_process_for_I_PCM_macroblocks();
        ret |= pic.Inter_prediction_process();
        if (m.transform_size_8x8_flag == 0) ret |= pic.transform_decoding_process_for_4x4_luma_residual_blocks_inter(0, 0, 8, wl, Y);
        else ret |= pic.transform_decoding_process_for_8x8_luma_residual_blocks_inter(0, 0, 8, wl, m.LumaLevel8x8, Y);
        ret |= pic.transform_decoding_process_for_chroma_samples_inter(1, wc, Cb);
        ret |= pic.transform_decoding_process_for_chroma_samples_inter(0, wc, Cr);
        return ret;
    }

    // Solves mvd and intra-mode flags of one MB with the reference's own predictors.
    void solve_mb(CH264PictureBase &pic, int addr, const h264g_mb &r) {
        CH264MacroBlock &m = pic.m_mbs[addr];
        if (r.mb_type == H264G_MB_I4x4) {
            for (int b = 0; b < 16; b++) {
                m.prev_intra4x4_pred_mode_flag[b] = 1;
                int pred = -1;
                pic.getIntra4x4PredMode(b, pred, 0);
                int t = r.intra_pred_modes[b];
                if (t == pred) m.prev_intra4x4_pred_mode_flag[b] = 1;
                else { m.prev_intra4x4_pred_mode_flag[b] = 0; m.rem_intra4x4_pred_mode[b] = t < pred ? t : t - 1; }
                m.Intra4x4PredMode[b] = (uint8_t)t;
            }
        } else if (r.mb_type == H264G_MB_I8x8) {
            for (int b = 0; b < 4; b++) {
                m.prev_intra8x8_pred_mode_flag[b] = 1;
                int pred = -1;
                pic.getIntra8x8PredMode(b, pred, 0);
                int t = r.intra_pred_modes[b];
                if (t == pred) m.prev_intra8x8_pred_mode_flag[b] = 1;
                else { m.prev_intra8x8_pred_mode_flag[b] = 0; m.rem_intra8x8_pred_mode[b] = t < pred ? t : t - 1; }
                m.Intra8x8PredMode[b] = (uint8_t)t;
            }
        } else if (r.mb_type == H264G_MB_P || r.mb_type == H264G_MB_B) {
            memset(m.mvd_l0, 0, sizeof(m.mvd_l0));
            memset(m.mvd_l1, 0, sizeof(m.mvd_l1));
            int np = r.part_shape == H264G_PART_16x16 ? 1 : (r.part_shape == H264G_PART_8x8 ? 4 : 2);
            for (int p = 0; p < np; p++) {
                if (r.flags & H264G_MBF_DIRECT(p)) continue;  // direct: no mvd, the reference derives it
                int ns = r.part_shape == H264G_PART_8x8 ? m.NumSubMbPart[p] : 1;
                for (int s = 0; s < ns; s++) {
                    clear_motion(m);
                    if (pic.Inter_prediction_process() != 0) fail("Inter_prediction_process failed while solving mvd");
                    int blk = part_block(r, p, s);
                    int q = (blk / 8) * 2 + (blk % 4) / 2;
                    for (int l = 0; l < 2; l++) {
                        if (r.ref_idx[l][q] < 0) continue;
                        int32_t (&mv)[4][4][2] = l == 0 ? m.m_MvL0 : m.m_MvL1;
                        int32_t (&mvd)[4][4][2] = l == 0 ? m.mvd_l0 : m.mvd_l1;
                        for (int c = 0; c < 2; c++) mvd[p][s][c] = r.mv[l][blk][c] - mv[p][s][c];
                    }
                }
            }
        }
        clear_motion(m);
    }

    void check_mb(CH264PictureBase &pic, int addr, const h264g_mb &r, int pic_idx) {
        if (r.mb_type != H264G_MB_P && r.mb_type != H264G_MB_B && r.mb_type != H264G_MB_P_SKIP) return;
        CH264MacroBlock &m = pic.m_mbs[addr];
        int np = r.mb_type == H264G_MB_P_SKIP ? 1 : (r.part_shape == H264G_PART_16x16 ? 1 : (r.part_shape == H264G_PART_8x8 ? 4 : 2));
        for (int p = 0; p < np && r.mb_type == H264G_MB_B; p++) {
            // a direct quadrant: the reference derived its indices and its four 4x4 vectors
            if (!(r.flags & H264G_MBF_DIRECT(p))) continue;
            for (int l = 0; l < 2; l++) {
                const int pf = l == 0 ? m.m_PredFlagL0[p] : m.m_PredFlagL1[p];
                const int ref = l == 0 ? m.m_RefIdxL0[p] : m.m_RefIdxL1[p];
                const bool used = r.ref_idx[l][p] >= 0;
                bool bad = pf != (used ? 1 : 0) || (used && ref != r.ref_idx[l][p]);
                for (int s = 0; s < 4 && used; s++) {
                    const int blk = ((p >> 1) * 2 + (s >> 1)) * 4 + (p & 1) * 2 + (s & 1);
                    int32_t (&mv)[4][4][2] = l == 0 ? m.m_MvL0 : m.m_MvL1;
                    bad |= mv[p][s][0] != r.mv[l][blk][0] || mv[p][s][1] != r.mv[l][blk][1];
                }
                if (bad) {
                    if (mismatches < 20)
                        fprintf(stderr, "direct mismatch pic %d mb %d quadrant %d list %d: ref pf %d idx %d, gen idx %d\n",
                                pic_idx, addr, p, l, pf, ref, r.ref_idx[l][p]);
                    mismatches++;
                }
            }
        }
        for (int p = 0; p < np; p++) {
            if (r.mb_type == H264G_MB_B && (r.flags & H264G_MBF_DIRECT(p))) continue;
            int ns = (r.mb_type != H264G_MB_P_SKIP && r.part_shape == H264G_PART_8x8) ? m.NumSubMbPart[p] : 1;
            for (int s = 0; s < ns; s++) {
                int blk = r.mb_type == H264G_MB_P_SKIP ? 0 : part_block(r, p, s);
                int q = (blk / 8) * 2 + (blk % 4) / 2;
                for (int l = 0; l < 2; l++) {
                    if (r.ref_idx[l][q] < 0) continue;
                    int32_t (&mv)[4][4][2] = l == 0 ? m.m_MvL0 : m.m_MvL1;
                    if (mv[p][s][0] != r.mv[l][blk][0] || mv[p][s][1] != r.mv[l][blk][1]) {
                        if (mismatches < 20)
                            fprintf(stderr, "MV mismatch pic %d mb %d part %d sub %d list %d: ref (%d,%d) gen (%d,%d)%s\n",
                                    pic_idx, addr, p, s, l, mv[p][s][0], mv[p][s][1], r.mv[l][blk][0], r.mv[l][blk][1],
                                    r.mb_type == H264G_MB_P_SKIP ? " [P_Skip]" : "");
                        mismatches++;
                    }
                }
            }
        }
    }

    void write_mvd(const CH264MacroBlock *ms, const h264g_mb *mbs, uint32_t nmb, int idx) {
        std::vector<int16_t> buf((size_t)nmb * 2 * 16 * 2, 0);
        for (uint32_t a = 0; a < nmb; a++) {
            const h264g_mb &r = mbs[a];
            if (r.mb_type != H264G_MB_P && r.mb_type != H264G_MB_B) continue;
            int np = r.part_shape == H264G_PART_16x16 ? 1 : (r.part_shape == H264G_PART_8x8 ? 4 : 2);
            for (int p = 0; p < np; p++) {
                if (r.flags & H264G_MBF_DIRECT(p)) continue;  // direct: no mvd
                int ns = 1;
                if (r.part_shape == H264G_PART_8x8)
                    ns = r.sub_part_shape[p] == H264G_SUB_8x8 ? 1 : (r.sub_part_shape[p] == H264G_SUB_4x4 ? 4 : 2);
                for (int s = 0; s < ns; s++) {
                    int blk = part_block(r, p, s);
                    int q = (blk / 8) * 2 + (blk % 4) / 2;
                    for (int l = 0; l < 2; l++) {
                        if (r.ref_idx[l][q] < 0) continue;
                        const int32_t (&md)[4][4][2] = l == 0 ? ms[a].mvd_l0 : ms[a].mvd_l1;
                        for (int c = 0; c < 2; c++) buf[(((size_t)a * 2 + l) * 16 + blk) * 2 + c] = (int16_t)md[p][s][c];
                    }
                }
            }
        }
        char path[1024];
        snprintf(path, sizeof(path), "%s/mvd_%03d.bin", out.c_str(), idx);
        FILE *f = fopen(path, "wb");
        if (!f) fail("cannot write output");
        fwrite(buf.data(), sizeof(int16_t), buf.size(), f);
        fclose(f);
        // the solved intra-mode syntax: uint8 [MB][16][2] (prev flag, rem) per
        // luma4x4BlkIdx of I4x4 MBs, per luma8x8BlkIdx (first four) of I8x8 MBs, else 0
        std::vector<uint8_t> im((size_t)nmb * 16 * 2, 0);
        for (uint32_t a = 0; a < nmb; a++) {
            uint8_t *o = &im[(size_t)a * 32];
            if (mbs[a].mb_type == H264G_MB_I4x4) {
                for (int b = 0; b < 16; b++) {
                    o[2 * b] = (uint8_t)ms[a].prev_intra4x4_pred_mode_flag[b];
                    o[2 * b + 1] = ms[a].prev_intra4x4_pred_mode_flag[b] ? 0 : (uint8_t)ms[a].rem_intra4x4_pred_mode[b];
                }
            } else if (mbs[a].mb_type == H264G_MB_I8x8) {
                for (int b = 0; b < 4; b++) {
                    o[2 * b] = (uint8_t)ms[a].prev_intra8x8_pred_mode_flag[b];
                    o[2 * b + 1] = ms[a].prev_intra8x8_pred_mode_flag[b] ? 0 : (uint8_t)ms[a].rem_intra8x8_pred_mode[b];
                }
            }
        }
        snprintf(path, sizeof(path), "%s/imode_%03d.bin", out.c_str(), idx);
        f = fopen(path, "wb");
        if (!f) fail("cannot write output");
        fwrite(im.data(), 1, im.size(), f);
        fclose(f);
    }

    void write_planes(CH264PictureBase &pic, const char *tag, int idx) {
        char path[1024];
        snprintf(path, sizeof(path), "%s/%s_%03d.yuv", out.c_str(), tag, idx);
        FILE *f = fopen(path, "wb");
        if (!f) fail("cannot write output");
        size_t n = (size_t)W * H * 256 * 3 / 2;
        fwrite(pic.m_pic_buff_luma, 1, n, f);  // Y, Cb, Cr are one contiguous block
        fclose(f);
        if (bgr && std::strcmp(tag, "pic") == 0) {
            const uint32_t w = (uint32_t)W * 16, h = (uint32_t)H * 16, stride = (w * 3 + 3) / 4 * 4;
            std::vector<uint8_t> buf((size_t)stride * h, 0);
            if (pic.convertYuv420pToBgr24(w, h, pic.m_pic_buff_luma, buf.data(), stride) != 0) fail("convertYuv420pToBgr24");
            snprintf(path, sizeof(path), "%s/bgr_%03d.raw", out.c_str(), idx);
            f = fopen(path, "wb");
            if (!f) fail("cannot write output");
            fwrite(buf.data(), 1, buf.size(), f);
            fclose(f);
        }
    }

    // Field macroblocks of MBAFF frames clamp the horizontal reference coordinate to the
    // field view's row pitch, twice the picture width (H264InterPrediction.cpp:2363,
    // 2504-2507 with H264PictureBase.cpp:209-210), so on the bottom field's last chroma
    // row the reference reads up to PicWidthInSamplesC - 1 bytes past the end of its
    // frame buffer.  main() pins glibc's mmap threshold so that frame buffers are page
    // mappings of their own; the bytes after a buffer up to the page end are then zero,
    // and this checks that they cover the read.
    void check_tail(const CH264PictureBase &pic) const {
        const size_t n = (size_t)W * H * 384, need = (size_t)W * 8;
        if (n < 256 * 1024) fail("MBAFF frame buffers below 256 KiB may come from the heap (glibc's top chunk)");
        // glibc marks a chunk it mapped with mmap by bit 1 of the size word before it
        const size_t chunk_size = ((const size_t *)pic.m_pic_buff_luma)[-1];
        if (!(chunk_size & 2)) fail("frame buffer came from the heap, not a page mapping of its own");
        const uint8_t *end = pic.m_pic_buff_luma + n;
        const size_t slack = (size_t)(-(uintptr_t)end & 4095);
        if (slack < need) fail("frame buffer ends too close to its page end for the field-MB overread");
        for (size_t i = 0; i < need; i++)
            if (end[i]) fail("bytes after the frame buffer are not zero");
    }

    void run() {
        W = gp.width_mbs; H = gp.height_mbs;
        setup_headers();
        int nslots = h264g_gen_dpb_slots(gen);
        slots.resize(nslots);
        memset(dpb, 0, sizeof(dpb));
        for (int s = 0; s < nslots; s++) { slots[s] = new CH264Picture; dpb[s] = slots[s]; }
        // Every DPB entry's frame buffer is allocated up front by the reference's own
        // init() (later inits keep it, H264PictureBase.cpp:154-157), before any other
        // large allocation, so that each is a page mapping of its own (check_tail).
        for (int s = 0; s < nslots; s++) {
            CH264PictureBase &f = slots[s]->m_picture_frame;
            f.m_picture_coded_type = H264_PICTURE_CODED_TYPE_FRAME;
            if (f.init(sh) != 0) fail("picture init");
            if (gp.mbaff) check_tail(f);
        }
        int npics = h264g_gen_num_pictures(gen);
        for (int i = 0; i < npics; i++) {
            h264g_pic_desc d;
            const h264g_mb *mbs;
            uint32_t nmb;
            const int16_t *lv;
            uint64_t nlv;
            h264g_gen_picture(gen, i, &d, &mbs, &nmb, &lv, &nlv);
            if ((int)nmb != W * H) fail("MB count");
            set_slice(d);
            CH264Picture *cur = slots[d.dst_slot];
            CH264PictureBase &pic = cur->m_picture_frame;
            // getOneEmptyPicture + reset (H264PictureBase.cpp:737-743), decode_one_slice init (H264Picture.cpp:124-160)
            cur->reset();
            pic.reset();
            cur->m_picture_top_filed.reset();
            cur->m_picture_bottom_filed.reset();
            cur->m_picture_coded_type = H264_PICTURE_CODED_TYPE_FRAME;
            pic.m_picture_coded_type = H264_PICTURE_CODED_TYPE_FRAME;
            pic.m_parent = cur;
            memcpy(pic.m_dpb, dpb, sizeof(dpb));
            cur->m_current_picture_ptr = &pic;
            if (pic.init(sh) != 0) fail("picture init");
            if (gp.mbaff) check_tail(pic);
            cur->m_picture_top_filed.m_picture_coded_type = H264_PICTURE_CODED_TYPE_TOP_FIELD;
            cur->m_picture_top_filed.m_parent = cur;
            cur->m_picture_top_filed.init(sh);
            cur->m_picture_bottom_filed.m_picture_coded_type = H264_PICTURE_CODED_TYPE_BOTTOM_FIELD;
            cur->m_picture_bottom_filed.m_parent = cur;
            cur->m_picture_bottom_filed.init(sh);
            pic.TopFieldOrderCnt = pic.BottomFieldOrderCnt = pic.PicOrderCnt = d.poc;
            cur->PicOrderCnt = d.poc;
            for (int l = 0; l < 2; l++) {
                CH264Picture **list = l == 0 ? pic.m_RefPicList0 : pic.m_RefPicList1;
                for (int r = 0; r < 16; r++) list[r] = nullptr;
                for (int r = 0; r < d.num_ref_idx_active[l] && r < 16; r++) list[r] = slots[d.ref_slot[l][r]];
            }
            pic.m_RefPicList0Length = d.num_ref_idx_active[0];
            pic.m_RefPicList1Length = d.num_ref_idx_active[1];
            pic.m_h264_slice_data.mb_field_decoding_flag = 0;
            pic.m_h264_slice_data.slice_number = 1;

            // pass 1: solve predicted syntax elements
            std::vector<DenseLevels> dl(nmb);
            int qp_prev = 0;
            for (uint32_t a = 0; a < nmb; a++) {
                expand_levels(mbs[a], lv, dl[a]);
                set_position(pic, (int)a, mbs[a]);
                fill_mb(pic, (int)a, mbs[a], dl[a], a == 0 ? mbs[a].qp_y : qp_prev);
                qp_prev = mbs[a].qp_y;
                solve_mb(pic, (int)a, mbs[a]);
                if (decode_mb(pic, (int)a, mbs[a]) != 0) fail("decode failed in pass 1");
            }
            // pass 2: clean decode with the solved syntax, timed
            // Each MB's parsed syntax is restored right before it is reconstructed, so
            // later MBs are all-zero (unparsed) exactly as in CH264SliceData::slice_data.
            std::vector<CH264MacroBlock> saved(pic.m_mbs, pic.m_mbs + nmb);
            for (uint32_t a = 0; a < nmb; a++) {
                CH264MacroBlock &m = saved[a];
                clear_motion(m);
                memset(m.Intra4x4PredMode, 0, sizeof(m.Intra4x4PredMode));
                memset(m.Intra8x8PredMode, 0, sizeof(m.Intra8x8PredMode));
            }
            memset(pic.m_pic_buff_luma, 0, (size_t)W * H * 384);
            memset((void *)pic.m_mbs, 0, sizeof(CH264MacroBlock) * nmb);
            auto t0 = std::chrono::steady_clock::now();
            for (uint32_t a = 0; a < nmb; a++) {
                pic.m_mbs[a] = saved[a];
                set_position(pic, (int)a, mbs[a]);
                if (decode_mb(pic, (int)a, mbs[a]) != 0) fail("decode failed in pass 2");
            }
            if (predeblock) {
                auto tp = std::chrono::steady_clock::now();
                recon_seconds += std::chrono::duration<double>(tp - t0).count();
                write_planes(pic, "pre", i);
                t0 = std::chrono::steady_clock::now();
            }
            pic.Deblocking_filter_process();
            auto t1 = std::chrono::steady_clock::now();
            recon_seconds += std::chrono::duration<double>(t1 - t0).count();
            for (uint32_t a = 0; a < nmb; a++) check_mb(pic, (int)a, mbs[a], i);
            write_planes(pic, "pic", i);
            if (mvd) write_mvd(saved.data(), mbs, nmb, i);
            // end of the picture (H264PictureBase.cpp:694-700) and the reference marking
            // state used by later pictures (a B picture's co-located picture: 8.4.1.2.1,
            // H264InterPrediction.cpp:1004-1011)
            pic.m_is_decode_finished = 1;
            cur->m_is_decode_finished = 1;
            cur->m_picture_coded_type_marked_as_refrence = H264_PICTURE_CODED_TYPE_FRAME;
            pic.reference_marked_type = H264_PICTURE_MARKED_AS_used_for_short_term_reference;
            cur->reference_marked_type = H264_PICTURE_MARKED_AS_used_for_short_term_reference;
        }
    }
};

}  // namespace

int main(int argc, char **argv) {
    // A fixed threshold (no dynamic adjustment): every frame buffer of 64 KiB or more is
    // a page mapping of its own (see Harness::check_tail).
    mallopt(M_MMAP_THRESHOLD, 64 * 1024);
    int config = 1;
    h264g_gen_params p;
    bool have = false;
    Harness hs;
    bool timing = false;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> const char * { if (i + 1 >= argc) fail("missing value"); return argv[++i]; };
        if (a == "--config") { config = atoi(next()); h264g_gen_default_params(config, &p); have = true; }
        else if (!have) fail("--config must come first");
        else if (a == "--frames") p.n_frames = atoi(next());
        else if (a == "--width") p.width_mbs = atoi(next());
        else if (a == "--height") p.height_mbs = atoi(next());
        else if (a == "--seed") p.seed = strtoull(next(), nullptr, 0);
        else if (a == "--ipcm") p.ipcm_permille = atoi(next());
        else if (a == "--wp") { const char *v = next(); sscanf(v, "%d,%d", &p.weighted_pred_flag, &p.weighted_bipred_idc); }
        else if (a == "--cqp") { const char *v = next(); sscanf(v, "%d,%d", &p.chroma_qp_index_offset, &p.second_chroma_qp_index_offset); }
        else if (a == "--deblock") p.disable_deblocking_filter_idc = atoi(next());
        else if (a == "--refs") p.max_refs = atoi(next());
        else if (a == "--bframes") p.b_frames = atoi(next());
        else if (a == "--t8x8") p.t8x8 = atoi(next());
        else if (a == "--mbaff") p.mbaff = atoi(next());
        else if (a == "--cip") p.constrained_intra_pred = atoi(next());
        else if (a == "--direct") p.direct_permille = atoi(next());
        else if (a == "--out") hs.out = next();
        else if (a == "--predeblock") hs.predeblock = true;
        else if (a == "--mvd") hs.mvd = true;
        else if (a == "--bgr") hs.bgr = true;
        else if (a == "--time") timing = true;
        else fail("unknown argument");
    }
    if (!have || hs.out.empty()) fail("usage: --config N ... --out DIR");
    hs.gp = p;
    if (h264g_gen_create(&p, &hs.gen) != 0) fail("generator");
    hs.run();
    uint64_t px = (uint64_t)p.width_mbs * p.height_mbs * 256 * h264g_gen_num_pictures(hs.gen);
    printf("{\"pictures\": %d, \"digest\": \"%016llx\", \"mv_mismatches\": %ld, \"recon_seconds\": %.6f, \"mpix\": %.6f}\n",
           h264g_gen_num_pictures(hs.gen), (unsigned long long)h264g_gen_digest(hs.gen), hs.mismatches,
           timing ? hs.recon_seconds : 0.0, px / 1e6);
    return hs.mismatches ? 3 : 0;
}
