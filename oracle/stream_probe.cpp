// stream_probe.cpp -- TEST INFRASTRUCTURE ONLY: a CPU probe of how the reference decoder
// handles field pictures (PAFF) and temporal direct prediction.  Nothing on the product
// path links or runs it.
//
// It writes a small Annex-B stream of its own (CAVLC, Main profile, frame_mbs_only_flag 0,
// no MBAFF): eight frames in the decode order of the 8-frame, 2-B-frame test sequence
// (I0 P3 B1 B2 P6 B4 B5 P7), then a top/bottom field pair (two P fields or two I
// fields, or none), then one more P frame so that the pair is closed by the decoder's
// normal end-of-picture path.  The B pictures' skipped MBs are B_Skip: spatial direct,
// or temporal direct with TEMPORAL.  I frames
// are all I_PCM with seeded samples, P and B pictures are one skip run.  The reference's
// own CH264VideoDecoder (all of its translation units, built unchanged) decodes the
// stream; the output callback reports, per decoded picture, the DPB entry it was decoded
// into (H264PicturesGOP.cpp:27-50 allocates 16 entries, getOneEmptyPicture
// H264PictureBase.cpp:668-687 picks the next one), the frame-view height the entry was
// initialised with (CH264PictureBase::init, H264PictureBase.cpp:128-233) and the size of
// the buffer behind it.  Built with AddressSanitizer (host code only), the run also shows
// whether a field decoded into a never-used entry writes past that entry's buffer.
//
// usage: stream_probe W H REORDER FIELDS [stream.h264] [TEMPORAL]
//   W, H      frame size in macroblocks (H even)
//   REORDER   max_num_reorder_frames written into the VUI, or -1 for no VUI (the
//             reference then infers it from level_idc, H264SPS.cpp:365-395)
//   FIELDS    P: the field pair is two P fields; I: two I fields; N: no field pair
//   TEMPORAL  1: direct_spatial_mv_pred_flag 0 in the B pictures (temporal direct);
//             2: the same with two active references per list in the B pictures
#include <malloc.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "H264VideoDecoder.h"

namespace {

struct BitWriter {
    std::vector<uint8_t> bytes;
    int nbits = 0;
    void u(int n, uint32_t v) {
        for (int i = n - 1; i >= 0; i--) {
            if (nbits % 8 == 0) bytes.push_back(0);
            if ((v >> i) & 1) bytes.back() |= (uint8_t)(0x80 >> (nbits % 8));
            nbits++;
        }
    }
    void ue(uint32_t v) {
        uint32_t x = v + 1;
        int len = 0;
        while ((x >> len) > 1) len++;
        u(len, 0);
        u(len + 1, x);
    }
    void se(int v) { ue(v > 0 ? 2 * v - 1 : -2 * v); }
    void align_zero() {
        while (nbits % 8) u(1, 0);
    }
    void trailing() {
        u(1, 1);
        align_zero();
    }
};

// Annex B start code + NAL header + payload with emulation-prevention bytes (7.4.1).
void put_nal(std::vector<uint8_t> &out, int ref_idc, int type, const std::vector<uint8_t> &rbsp) {
    const uint8_t sc[4] = {0, 0, 0, 1};
    out.insert(out.end(), sc, sc + 4);
    out.push_back((uint8_t)((ref_idc << 5) | type));
    int zeros = 0;
    for (uint8_t b : rbsp) {
        if (zeros >= 2 && b <= 3) {
            out.push_back(3);
            zeros = 0;
        }
        out.push_back(b);
        zeros = b == 0 ? zeros + 1 : 0;
    }
}

struct Pic {
    char kind;      // 'I', 'P', 'B'
    int ref;        // nal_ref_idc != 0
    int idr;
    int frame_num;
    int poc_lsb;
    int field;      // 0 frame, 1 top field, 2 bottom field
};

uint32_t g_rng = 12345;
uint8_t next_sample() {
    g_rng = g_rng * 1103515245u + 12345u;
    return (uint8_t)(16 + ((g_rng >> 16) % 220));  // 16..235: no zero bytes
}

std::vector<uint8_t> make_stream(int W, int H, int reorder, char fields, int temporal) {
    std::vector<uint8_t> out;
    {   // seq_parameter_set_rbsp (7.3.2.1.1)
        BitWriter b;
        b.u(8, 77);      // Main
        b.u(8, 0);       // constraint flags + reserved
        b.u(8, 40);      // level 4.0
        b.ue(0);         // seq_parameter_set_id
        b.ue(0);         // log2_max_frame_num_minus4 -> MaxFrameNum 16
        b.ue(0);         // pic_order_cnt_type
        b.ue(2);         // log2_max_pic_order_cnt_lsb_minus4 -> 64
        b.ue(4);         // max_num_ref_frames
        b.u(1, 0);       // gaps_in_frame_num_value_allowed_flag
        b.ue(W - 1);     // pic_width_in_mbs_minus1
        b.ue(H / 2 - 1); // pic_height_in_map_units_minus1 (map units are MB pairs)
        b.u(1, 0);       // frame_mbs_only_flag
        b.u(1, 0);       // mb_adaptive_frame_field_flag
        b.u(1, 1);       // direct_8x8_inference_flag
        b.u(1, 0);       // frame_cropping_flag
        b.u(1, reorder >= 0 ? 1 : 0);  // vui_parameters_present_flag
        if (reorder >= 0) {  // vui_parameters (E.1.1) with bitstream_restriction only
            for (int i = 0; i < 4; i++) b.u(1, 0);  // aspect, overscan, video_signal, chroma_loc
            b.u(1, 0);       // timing_info_present_flag
            b.u(1, 0);       // nal_hrd_parameters_present_flag
            b.u(1, 0);       // vcl_hrd_parameters_present_flag
            b.u(1, 0);       // pic_struct_present_flag
            b.u(1, 1);       // bitstream_restriction_flag
            b.u(1, 1);       // motion_vectors_over_pic_boundaries_flag
            b.ue(0);         // max_bytes_per_pic_denom
            b.ue(0);         // max_bits_per_mb_denom
            b.ue(16);        // log2_max_mv_length_horizontal
            b.ue(16);        // log2_max_mv_length_vertical
            b.ue(reorder);   // max_num_reorder_frames
            b.ue(4);         // max_dec_frame_buffering
        }
        b.trailing();
        put_nal(out, 3, 7, b.bytes);
    }
    {   // pic_parameter_set_rbsp (7.3.2.2)
        BitWriter b;
        b.ue(0); b.ue(0);   // pps id, sps id
        b.u(1, 0);          // entropy_coding_mode_flag (CAVLC)
        b.u(1, 0);          // bottom_field_pic_order_in_frame_present_flag
        b.ue(0);            // num_slice_groups_minus1
        b.ue(0); b.ue(0);   // num_ref_idx_l0/l1_default_active_minus1
        b.u(1, 0);          // weighted_pred_flag
        b.u(2, 0);          // weighted_bipred_idc
        b.se(0); b.se(0);   // pic_init_qp_minus26, pic_init_qs_minus26
        b.se(0);            // chroma_qp_index_offset
        b.u(1, 1);          // deblocking_filter_control_present_flag
        b.u(1, 0);          // constrained_intra_pred_flag
        b.u(1, 0);          // redundant_pic_cnt_present_flag
        b.trailing();
        put_nal(out, 3, 8, b.bytes);
    }
    // decode order of the 8-frame sequence with two B frames between anchors, then a
    // top-field-first P field pair and a closing P frame
    const std::vector<Pic> pics = {
        {'I', 1, 1, 0, 0, 0},  {'P', 1, 0, 1, 6, 0},  {'B', 0, 0, 2, 2, 0},  {'B', 0, 0, 2, 4, 0},
        {'P', 1, 0, 2, 12, 0}, {'B', 0, 0, 3, 8, 0},  {'B', 0, 0, 3, 10, 0}, {'P', 1, 0, 3, 14, 0},
        {fields, 1, 0, 4, 16, 1}, {fields, 1, 0, 4, 17, 2}, {'P', 1, 0, 5, 18, 0},
    };
    for (const Pic &p : pics) {
        if (p.kind == 'N') continue;
        BitWriter b;
        const int mbs = W * H / (p.field ? 2 : 1);
        b.ue(0);  // first_mb_in_slice
        b.ue(p.kind == 'I' ? 7 : (p.kind == 'P' ? 5 : 6));
        b.ue(0);  // pic_parameter_set_id
        b.u(4, (uint32_t)p.frame_num);
        b.u(1, p.field ? 1 : 0);             // field_pic_flag
        if (p.field) b.u(1, p.field == 2);   // bottom_field_flag
        if (p.idr) b.ue(0);                  // idr_pic_id
        b.u(6, (uint32_t)p.poc_lsb);
        if (p.kind == 'B') b.u(1, temporal ? 0 : 1);  // direct_spatial_mv_pred_flag
        if (p.kind == 'B' && temporal == 2) {  // num_ref_idx_active_override_flag: 2 + 2
            b.u(1, 1);
            b.ue(1);
            b.ue(1);
        } else if (p.kind != 'I') {
            b.u(1, 0);                       // num_ref_idx_active_override_flag
        }
        if (p.kind != 'I') b.u(1, 0);        // ref_pic_list_modification_flag_l0
        if (p.kind == 'B') b.u(1, 0);        // ref_pic_list_modification_flag_l1
        if (p.ref) {                          // dec_ref_pic_marking
            if (p.idr) { b.u(1, 0); b.u(1, 0); }
            else b.u(1, 0);                   // adaptive_ref_pic_marking_mode_flag
        }
        b.se(0);   // slice_qp_delta
        b.ue(0);   // disable_deblocking_filter_idc
        b.se(0); b.se(0);
        if (p.kind == 'I') {
            for (int m = 0; m < mbs; m++) {
                b.ue(25);  // I_PCM
                b.align_zero();
                for (int k = 0; k < 384; k++) b.u(8, next_sample());
            }
        } else {
            b.ue((uint32_t)mbs);  // one skip run over the whole picture
        }
        b.trailing();
        put_nal(out, p.ref ? 2 : 0, p.idr ? 5 : 1, b.bytes);
    }
    return out;
}

struct Seen {
    int entry;
    int num;          // m_PicNumCnt: decode order
    int coded_type;
    int frame_rows;   // luma rows of the frame view the entry was initialised with
    size_t buf_bytes; // size of the allocation behind the entry's frame view
    int poc;
    long nonzero_even, nonzero_odd;  // non-zero luma samples on even / odd rows (within the buffer)
};

std::vector<Seen> g_seen;
int g_W = 0, g_H = 0;

int on_picture(CH264Picture *pic, void *, int) {
    if (!pic) return -1;
    CH264PictureBase &f = pic->m_picture_frame;
    Seen s{};
    s.entry = -1;
    for (int i = 0; i < 16; i++)
        if (f.m_dpb[i] == pic) s.entry = i;
    s.num = f.m_PicNumCnt;
    s.coded_type = (int)pic->m_picture_coded_type;
    s.frame_rows = f.PicHeightInSamplesL;
    s.buf_bytes = f.m_pic_buff_luma ? malloc_usable_size(f.m_pic_buff_luma) : 0;
    s.poc = f.PicOrderCnt;
    const size_t pitch = (size_t)g_W * 16;
    const size_t rows = std::min<size_t>((size_t)g_H * 16, s.buf_bytes / pitch);
    for (size_t y = 0; y < rows; y++)
        for (size_t x = 0; x < pitch; x++)
            if (f.m_pic_buff_luma[y * pitch + x]) (y & 1 ? s.nonzero_odd : s.nonzero_even)++;
    g_seen.push_back(s);
    return 0;
}

const char *type_name(int t) {
    switch (t) {
    case H264_PICTURE_CODED_TYPE_FRAME: return "frame";
    case H264_PICTURE_CODED_TYPE_COMPLEMENTARY_FIELD_PAIR: return "field";
    default: return "?";
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s W H REORDER FIELDS [stream.h264] [TEMPORAL]\n", argv[0]);
        return 2;
    }
    setvbuf(stdout, nullptr, _IOLBF, 0);  // keep the decoder's log in order with sanitizer reports
    g_W = atoi(argv[1]);
    g_H = atoi(argv[2]);
    const int reorder = atoi(argv[3]);
    const char fields = argv[4][0];
    const char *path = argc > 5 ? argv[5] : "/tmp/stream_probe.h264";
    const int temporal = argc > 6 ? atoi(argv[6]) : 0;
    if (g_W <= 0 || g_H <= 0 || g_H % 2 || (fields != 'P' && fields != 'I' && fields != 'N')) return 2;
    std::vector<uint8_t> s = make_stream(g_W, g_H, reorder, fields, temporal);
    FILE *fp = fopen(path, "wb");
    if (!fp || fwrite(s.data(), 1, s.size(), fp) != s.size()) return 3;
    fclose(fp);

    CH264VideoDecoder dec;
    dec.set_output_frame_callback_functuin(on_picture, nullptr);
    int rc = dec.open(path);

    std::sort(g_seen.begin(), g_seen.end(), [](const Seen &a, const Seen &b) { return a.num < b.num; });
    const size_t frame_bytes = (size_t)g_W * g_H * 384;
    std::map<int, int> first_use;
    printf("STREAM_PROBE W=%d H=%d reorder=%d fields=%c temporal=%d decoder_rc=%d frame_buffer_bytes=%zu\n", g_W, g_H,
           reorder, fields, temporal, rc, frame_bytes);
    printf("%-4s %-6s %-5s %-10s %-10s %-12s %-5s %-12s %-12s\n", "num", "type", "entry", "reused", "frame_rows",
           "buffer", "poc", "nz_even_rows", "nz_odd_rows");
    for (const Seen &e : g_seen) {
        const bool reused = first_use.count(e.entry) != 0;
        if (!reused) first_use[e.entry] = e.num;
        printf("%-4d %-6s %-5d %-10s %-10d %-12zu %-5d %-12ld %-12ld\n", e.num, type_name(e.coded_type), e.entry,
               reused ? "yes" : "first-use", e.frame_rows, e.buf_bytes, e.poc, e.nonzero_even, e.nonzero_odd);
    }
    return 0;
}
