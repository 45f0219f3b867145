// Annex B byte-stream front end for frame pictures coded with CAVLC or CABAC: NAL units,
// SPS / PPS, slice headers, slice data (macroblock_layer, mb_pred / sub_mb_pred,
// residual_block_cavlc / residual_block_cabac)
// and the decoded picture buffer (POC, reference marking, reference picture lists, DPB
// slots), producing the h264g_pic_desc / h264g_mb / level records the GPU stages consume.
// It replaces the reference's parsing and picture management on the path to
// reconstruction: CH264NalUnit / CH264SPS / CH264PPS / CH264SliceHeader
// (H264NalUnit.cpp:122, H264SPS.cpp, H264PPS.cpp, H264SliceHeader.cpp:289-509),
// CH264SliceData::slice_data and CH264MacroBlock::macroblock_layer
// (H264SliceData.cpp:386-487, H264MacroBlock.cpp:912-1139), CH264ResidualBlockCavlc
// (H264ResidualBlockCavlc.cpp), CH264Cabac (H264Cabac.cpp) and the reference list /
// marking code (H264RefPicList.cpp).
// Motion vectors and intra prediction modes are derived by the motion field
// (h264g_mvfield_*, mvpred.cpp) from the parsed mvd and prev / rem syntax.
//
// Pictures are gathered in stream order on the calling thread (parameter sets, slice
// headers, each slice's RBSP); the slice data of the pictures after the one handed out is
// parsed ahead, one worker thread per picture.  A picture without B slices is derived
// there too (vectors and modes, in a motion field of its own); a B picture's records wait
// for the calling thread, which derives them in decoding order, after the co-located
// picture (entropy decoding reads only parsed syntax of its neighbours, never derived
// vectors).  The output is that of the in-order parse.
//
// Pictures may have several slices (in order, sharing one reference list, weight table
// and deblocking setting); each MB record carries its slice number, which makes
// neighbours in other slices unavailable.  SPS / PPS scaling matrices become the
// picture's lists as the reference derives them.  Refused with H264G_E_ARG (none occur
// in the reference's progressive sample streams): field pictures and MBAFF, slice groups,
// arbitrary slice order, chroma formats other than 4:2:0, bit depths above 8, lossless
// coding, SP / SI slices, temporal direct prediction.
#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "h264g.h"

extern "C" int32_t h264g_set_error(int32_t code, const char *fmt, ...);

namespace {

// ------------------------------------------------------------------ bit reader
struct Bits {
    const uint8_t *d = nullptr;
    size_t nbits = 0, pos = 0;
    bool over = false;  // read past the end
    uint32_t bit() {
        if (pos >= nbits) { over = true; return 0; }
        const uint32_t b = (d[pos >> 3] >> (7 - (pos & 7))) & 1u;
        pos++;
        return b;
    }
    uint32_t u(int n) {
        if (n <= 0) return 0;
        const uint32_t v = peek(n);
        skip(n);
        return v;
    }
    // the next n (1..32) bits without consuming them, zeros past the end (nbits is a
    // whole number of bytes): five bytes from the current one, shifted into place
    uint32_t peek(int n) const {
        const size_t nb = nbits >> 3, b0 = pos >> 3;
        uint64_t v = 0;
        for (size_t i = 0; i < 5; i++) v = (v << 8) | (b0 + i < nb ? d[b0 + i] : 0u);
        return (uint32_t)(((v << (pos & 7)) & 0xFFFFFFFFFFull) >> (40 - n));
    }
    void skip(int n) {  // past the end: flagged, and the position stays at the end
        pos += n;
        if (pos > nbits) { over = true; pos = nbits; }
    }
    uint32_t ue() {  // 9.1: leading zeros, a 1, as many more bits
        const uint32_t w = peek(32);
        if (!w) { skip(32); over = true; return 0; }  // more than 31 leading zeros
        const int z = __builtin_clz(w);
        skip(z + 1);
        return z ? ((1u << z) - 1u) + u(z) : 0u;
    }
    int32_t se() {  // 9.1.1
        const uint32_t k = ue();
        return (k & 1u) ? (int32_t)((k + 1) >> 1) : -(int32_t)(k >> 1);
    }
    int32_t te(int range) { return range > 1 ? (int32_t)ue() : (int32_t)!bit(); }  // 9.1.2 (range >= 1)
    bool aligned() const { return (pos & 7) == 0; }
    // more_rbsp_data(): anything before the rbsp_stop_one_bit (the last 1 bit)
    bool more_rbsp_data() const {
        if (pos >= nbits) return false;
        size_t last = nbits;
        while (last > pos) {
            const size_t p = last - 1;
            if ((d[p >> 3] >> (7 - (p & 7))) & 1u) break;
            last--;
        }
        return last > 0 && pos < last - 1;
    }
};

// NAL payload without its header byte and emulation-prevention bytes (7.3.1, 7.4.1)
void unescape(const uint8_t *p, size_t n, std::vector<uint8_t> &out) {
    out.clear();
    out.reserve(n);
    int zeros = 0;
    for (size_t i = 0; i < n; i++) {
        if (zeros >= 2 && p[i] == 3) { zeros = 0; continue; }
        out.push_back(p[i]);
        zeros = p[i] == 0 ? zeros + 1 : 0;
    }
}

// ------------------------------------------------------------------ CAVLC code tables
// The standard's code tables, each entry a codeword as written in the standard.
struct Code {
    const char *bits;
    int8_t a, b;  // coeff_token: TrailingOnes, TotalCoeff; total_zeros / run_before: value
};

// Table 9-5, 0 <= nC < 2
const Code kCoeffToken0[] = {
    {"1", 0, 0}, {"000101", 0, 1}, {"01", 1, 1}, {"00000111", 0, 2}, {"000100", 1, 2}, {"001", 2, 2},
    {"000000111", 0, 3}, {"00000110", 1, 3}, {"0000101", 2, 3}, {"00011", 3, 3},
    {"0000000111", 0, 4}, {"000000110", 1, 4}, {"00000101", 2, 4}, {"000011", 3, 4},
    {"00000000111", 0, 5}, {"0000000110", 1, 5}, {"000000101", 2, 5}, {"0000100", 3, 5},
    {"0000000001111", 0, 6}, {"00000000110", 1, 6}, {"0000000101", 2, 6}, {"00000100", 3, 6},
    {"0000000001011", 0, 7}, {"0000000001110", 1, 7}, {"00000000101", 2, 7}, {"000000100", 3, 7},
    {"0000000001000", 0, 8}, {"0000000001010", 1, 8}, {"0000000001101", 2, 8}, {"0000000100", 3, 8},
    {"00000000001111", 0, 9}, {"00000000001110", 1, 9}, {"0000000001001", 2, 9}, {"00000000100", 3, 9},
    {"00000000001011", 0, 10}, {"00000000001010", 1, 10}, {"00000000001101", 2, 10}, {"0000000001100", 3, 10},
    {"000000000001111", 0, 11}, {"000000000001110", 1, 11}, {"00000000001001", 2, 11}, {"00000000001100", 3, 11},
    {"000000000001011", 0, 12}, {"000000000001010", 1, 12}, {"000000000001101", 2, 12}, {"00000000001000", 3, 12},
    {"0000000000001111", 0, 13}, {"000000000000001", 1, 13}, {"000000000001001", 2, 13}, {"000000000001100", 3, 13},
    {"0000000000001011", 0, 14}, {"0000000000001110", 1, 14}, {"0000000000001101", 2, 14}, {"000000000001000", 3, 14},
    {"0000000000000111", 0, 15}, {"0000000000001010", 1, 15}, {"0000000000001001", 2, 15}, {"0000000000001100", 3, 15},
    {"0000000000000100", 0, 16}, {"0000000000000110", 1, 16}, {"0000000000000101", 2, 16}, {"0000000000001000", 3, 16},
};
// Table 9-5, 2 <= nC < 4
const Code kCoeffToken1[] = {
    {"11", 0, 0}, {"001011", 0, 1}, {"10", 1, 1}, {"000111", 0, 2}, {"00111", 1, 2}, {"011", 2, 2},
    {"0000111", 0, 3}, {"001010", 1, 3}, {"001001", 2, 3}, {"0101", 3, 3},
    {"00000111", 0, 4}, {"000110", 1, 4}, {"000101", 2, 4}, {"0100", 3, 4},
    {"00000100", 0, 5}, {"0000110", 1, 5}, {"0000101", 2, 5}, {"00110", 3, 5},
    {"000000111", 0, 6}, {"00000110", 1, 6}, {"00000101", 2, 6}, {"001000", 3, 6},
    {"00000001111", 0, 7}, {"000000110", 1, 7}, {"000000101", 2, 7}, {"000100", 3, 7},
    {"00000001011", 0, 8}, {"00000001110", 1, 8}, {"00000001101", 2, 8}, {"0000100", 3, 8},
    {"000000001111", 0, 9}, {"00000001010", 1, 9}, {"00000001001", 2, 9}, {"000000100", 3, 9},
    {"000000001011", 0, 10}, {"000000001110", 1, 10}, {"000000001101", 2, 10}, {"00000001100", 3, 10},
    {"000000001000", 0, 11}, {"000000001010", 1, 11}, {"000000001001", 2, 11}, {"00000001000", 3, 11},
    {"0000000001111", 0, 12}, {"0000000001110", 1, 12}, {"0000000001101", 2, 12}, {"000000001100", 3, 12},
    {"0000000001011", 0, 13}, {"0000000001010", 1, 13}, {"0000000001001", 2, 13}, {"0000000001100", 3, 13},
    {"0000000000111", 0, 14}, {"00000000001011", 1, 14}, {"0000000000110", 2, 14}, {"0000000001000", 3, 14},
    {"00000000001001", 0, 15}, {"00000000001000", 1, 15}, {"00000000001010", 2, 15}, {"0000000000001", 3, 15},
    {"00000000000111", 0, 16}, {"00000000000110", 1, 16}, {"00000000000101", 2, 16}, {"00000000000100", 3, 16},
};
// Table 9-5, 4 <= nC < 8
const Code kCoeffToken2[] = {
    {"1111", 0, 0}, {"001111", 0, 1}, {"1110", 1, 1}, {"001011", 0, 2}, {"01111", 1, 2}, {"1101", 2, 2},
    {"001000", 0, 3}, {"01100", 1, 3}, {"01110", 2, 3}, {"1100", 3, 3},
    {"0001111", 0, 4}, {"01010", 1, 4}, {"01011", 2, 4}, {"1011", 3, 4},
    {"0001011", 0, 5}, {"01000", 1, 5}, {"01001", 2, 5}, {"1010", 3, 5},
    {"0001001", 0, 6}, {"001110", 1, 6}, {"001101", 2, 6}, {"1001", 3, 6},
    {"0001000", 0, 7}, {"001010", 1, 7}, {"001001", 2, 7}, {"1000", 3, 7},
    {"00001111", 0, 8}, {"0001110", 1, 8}, {"0001101", 2, 8}, {"01101", 3, 8},
    {"00001011", 0, 9}, {"00001110", 1, 9}, {"0001010", 2, 9}, {"001100", 3, 9},
    {"000001111", 0, 10}, {"00001010", 1, 10}, {"00001101", 2, 10}, {"0001100", 3, 10},
    {"000001011", 0, 11}, {"000001110", 1, 11}, {"00001001", 2, 11}, {"00001100", 3, 11},
    {"000001000", 0, 12}, {"000001010", 1, 12}, {"000001101", 2, 12}, {"00001000", 3, 12},
    {"0000001101", 0, 13}, {"000000111", 1, 13}, {"000001001", 2, 13}, {"000001100", 3, 13},
    {"0000001001", 0, 14}, {"0000001100", 1, 14}, {"0000001011", 2, 14}, {"0000001010", 3, 14},
    {"0000000101", 0, 15}, {"0000001000", 1, 15}, {"0000000111", 2, 15}, {"0000000110", 3, 15},
    {"0000000001", 0, 16}, {"0000000100", 1, 16}, {"0000000011", 2, 16}, {"0000000010", 3, 16},
};
// Table 9-5, nC == -1 (chroma DC, 4:2:0)
const Code kCoeffTokenChromaDc[] = {
    {"01", 0, 0}, {"000111", 0, 1}, {"1", 1, 1}, {"000100", 0, 2}, {"000110", 1, 2}, {"001", 2, 2},
    {"000011", 0, 3}, {"0000011", 1, 3}, {"0000010", 2, 3}, {"000101", 3, 3},
    {"000010", 0, 4}, {"00000011", 1, 4}, {"00000010", 2, 4}, {"0000000", 3, 4},
};
// Tables 9-7 / 9-8: total_zeros of 4x4 blocks, tzVlcIndex 1..15 (one string per value)
const char *const kTotalZeros[15][16] = {
    {"1", "011", "010", "0011", "0010", "00011", "00010", "000011", "000010", "0000011", "0000010", "00000011",
     "00000010", "000000011", "000000010", "000000001"},
    {"111", "110", "101", "100", "011", "0101", "0100", "0011", "0010", "00011", "00010", "000011", "000010",
     "000001", "000000"},
    {"0101", "111", "110", "101", "0100", "0011", "100", "011", "0010", "00011", "00010", "000001", "00001",
     "000000"},
    {"00011", "111", "0101", "0100", "110", "101", "100", "0011", "011", "0010", "00010", "00001", "00000"},
    {"0101", "0100", "0011", "111", "110", "101", "100", "011", "0010", "00001", "0001", "00000"},
    {"000001", "00001", "111", "110", "101", "100", "011", "010", "0001", "001", "000000"},
    {"000001", "00001", "101", "100", "011", "11", "010", "0001", "001", "000000"},
    {"000001", "0001", "00001", "011", "11", "10", "010", "001", "000000"},
    {"000001", "000000", "0001", "11", "10", "001", "01", "00001"},
    {"00001", "00000", "001", "11", "10", "01", "0001"},
    {"0000", "0001", "001", "010", "1", "011"},
    {"0000", "0001", "01", "1", "001"},
    {"000", "001", "1", "01"},
    {"00", "01", "1"},
    {"0", "1"},
};
// Table 9-9a: total_zeros of chroma DC (4:2:0), tzVlcIndex 1..3
const char *const kTotalZerosChromaDc[3][4] = {{"1", "01", "001", "000"}, {"1", "01", "00"}, {"1", "0"}};
// Table 9-10: run_before, zerosLeft 1..6 and > 6
const char *const kRunBefore[7][15] = {
    {"1", "0"},
    {"1", "01", "00"},
    {"11", "10", "01", "00"},
    {"11", "10", "01", "001", "000"},
    {"11", "10", "011", "010", "001", "000"},
    {"11", "000", "001", "011", "010", "101", "100"},
    {"111", "110", "101", "100", "011", "010", "001", "0001", "00001", "000001", "0000001", "00000001", "000000001",
     "0000000001", "00000000001"},
};
// Table 9-4 (ChromaArrayType 1 or 2): codeNum of coded_block_pattern -> value, for
// Intra_4x4 / Intra_8x8 MBs and for inter MBs
const uint8_t kCbpIntra[48] = {47, 31, 15, 0,  23, 27, 29, 30, 7,  11, 13, 14, 39, 43, 45, 46,
                               16, 3,  5,  10, 12, 19, 21, 26, 28, 35, 37, 42, 44, 1,  2,  4,
                               8,  17, 18, 20, 24, 6,  9,  22, 25, 32, 33, 34, 36, 40, 38, 41};
const uint8_t kCbpInter[48] = {0,  16, 1,  2,  4,  8,  32, 3,  5,  10, 12, 15, 47, 7,  11, 13,
                               14, 6,  9,  31, 35, 37, 42, 44, 33, 34, 36, 40, 39, 43, 45, 46,
                               17, 18, 20, 24, 19, 21, 26, 28, 23, 27, 29, 30, 22, 25, 38, 41};

// A codeword table compiled for lookup.  The codes (at most 16 bits, prefix-free) are
// grouped by their number z of leading zeros; within a group an entry is found by the
// m(z) bits after the first 1, m(z) the longest such tail in the group (shorter codes fill
// every completion of their tail).  A code of zeros only, when the table has one, is
// matched first.
struct Vlc {
    struct E { uint8_t len; int8_t a, b; };  // len 0: no codeword
    struct Code { uint32_t code; int len, a, b; };
    std::vector<Code> codes;
    int8_t m[17] = {};
    uint16_t base[18] = {};
    std::vector<E> lut;
    E zero{0, 0, 0};
    void add(const char *s, int a, int b) {
        uint32_t c = 0;
        int n = 0;
        for (; s[n]; n++) c = (c << 1) | (uint32_t)(s[n] == '1');
        codes.push_back({c, n, a, b});
    }
    static int zeros_of(const Code &x) { return x.len - (32 - __builtin_clz(x.code)); }
    void build() {
        for (const Code &x : codes)
            if (x.code) m[zeros_of(x)] = (int8_t)std::max<int>(m[zeros_of(x)], x.len - zeros_of(x) - 1);
        for (int z = 0; z < 17; z++) base[z + 1] = (uint16_t)(base[z] + (1 << m[z]));
        lut.assign(base[17], E{0, 0, 0});
        for (const Code &x : codes) {
            if (!x.code) {
                zero = E{(uint8_t)x.len, (int8_t)x.a, (int8_t)x.b};
                continue;
            }
            const int z = zeros_of(x), tail = x.len - z - 1, fill = m[z] - tail;
            const uint32_t bits = x.code & ((1u << tail) - 1u);
            for (uint32_t k = 0; k < (1u << fill); k++)
                lut[base[z] + (bits << fill) + k] = E{(uint8_t)x.len, (int8_t)x.a, (int8_t)x.b};
        }
    }
    // the entry whose codeword is next in the stream (consumed), or null
    const E *read(Bits &bs) const {
        const uint32_t w = bs.peek(16);
        if (zero.len && (w >> (16 - zero.len)) == 0) {
            bs.skip(zero.len);
            return &zero;
        }
        if (!w) return nullptr;
        const int z = __builtin_clz(w) - 16;
        const E &x = lut[base[z] + (((w << (z + 1)) & 0xFFFFu) >> (16 - m[z]))];
        if (!x.len) return nullptr;
        bs.skip(x.len);
        return &x;
    }
};

struct Tables {
    Vlc coeff_token[4];  // nC 0-1, 2-3, 4-7, chroma DC
    Vlc total_zeros[15], total_zeros_cdc[3], run_before[7];
    Tables() {
        for (const Code &c : kCoeffToken0) coeff_token[0].add(c.bits, c.a, c.b);
        for (const Code &c : kCoeffToken1) coeff_token[1].add(c.bits, c.a, c.b);
        for (const Code &c : kCoeffToken2) coeff_token[2].add(c.bits, c.a, c.b);
        for (const Code &c : kCoeffTokenChromaDc) coeff_token[3].add(c.bits, c.a, c.b);
        for (int t = 0; t < 15; t++)
            for (int v = 0; v < 16 && kTotalZeros[t][v]; v++) total_zeros[t].add(kTotalZeros[t][v], v, 0);
        for (int t = 0; t < 3; t++)
            for (int v = 0; v < 4 && kTotalZerosChromaDc[t][v]; v++) total_zeros_cdc[t].add(kTotalZerosChromaDc[t][v], v, 0);
        for (int t = 0; t < 7; t++)
            for (int v = 0; v < 15 && kRunBefore[t][v]; v++) run_before[t].add(kRunBefore[t][v], v, 0);
        for (Vlc &v : coeff_token) v.build();
        for (Vlc &v : total_zeros) v.build();
        for (Vlc &v : total_zeros_cdc) v.build();
        for (Vlc &v : run_before) v.build();
    }
};
const Tables &tables() {
    static const Tables t;
    return t;
}

// ------------------------------------------------------------------ parameter sets
struct ScalingLists {
    bool present_matrix = false;                 // seq_ / pic_scaling_matrix_present_flag
    uint8_t present[8] = {}, use_default[8] = {};  // per list: *_scaling_list_present_flag, useDefault
    uint8_t l4[6][16] = {}, l8[2][64] = {};      // as parsed, scan order; zero when not sent
};
struct Sps {
    bool valid = false;
    int profile_idc = 0, constraint3 = 0, level_idc = 0;
    int chroma_format_idc = 1, log2_max_frame_num = 4, poc_type = 0, log2_max_poc_lsb = 4;
    // pic_order_cnt_type 1 (7.4.2.1.1)
    int delta_poc_always_zero = 0, offset_non_ref = 0, offset_top_to_bottom = 0, n_cycle = 0;
    int offset_ref[256] = {}, expected_delta_cycle = 0;
    int max_num_ref_frames = 0, gaps = 0, width_mbs = 0, height_map_units = 0, frame_mbs_only = 1;
    int direct_8x8_inference = 0;
    int max_num_reorder_frames = -1;  // VUI bitstream_restriction, -1 when absent
    int max_dec_frame_buffering = -1;
    ScalingLists sl;
};
struct Pps {
    bool valid = false;
    ScalingLists sl;
    int sps_id = 0, cabac = 0, bottom_field_pic_order = 0, num_slice_groups = 1;
    int num_ref_idx_default[2] = {1, 1}, weighted_pred = 0, weighted_bipred_idc = 0, pic_init_qp = 26;
    int chroma_qp_index_offset = 0, second_chroma_qp_index_offset = 0, deblocking_filter_control = 0;
    int constrained_intra_pred = 0, redundant_pic_cnt_present = 0, transform_8x8_mode = 0;
};

// scaling_list() (7.3.2.1.1.1) as the reference parses it (H264CommonFunc.cpp:48-67): every
// entry is stored, useDefaultScalingMatrixFlag when the first delta makes nextScale 0.
int parse_scaling_list(Bits &bs, uint8_t *list, int n, uint8_t &use_default) {
    int last = 8, next = 8;
    for (int j = 0; j < n; j++) {
        if (next != 0) {
            const int delta = bs.se();
            if (delta < -128 || delta > 127) return h264g_set_error(H264G_E_ARG, "delta_scale %d", delta);
            next = (last + delta + 256) % 256;
            use_default = j == 0 && next == 0;
        }
        list[j] = (uint8_t)(next == 0 ? last : next);
        last = list[j];
    }
    return 0;
}
// The scaling_list_present flags and lists of an SPS or PPS (4:2:0: six 4x4, two 8x8).
int parse_scaling_lists(Bits &bs, ScalingLists &sl, int n) {
    sl.present_matrix = true;
    for (int i = 0; i < n; i++) {
        sl.present[i] = (uint8_t)bs.u(1);
        if (!sl.present[i]) continue;
        if (int rc = i < 6 ? parse_scaling_list(bs, sl.l4[i], 16, sl.use_default[i])
                           : parse_scaling_list(bs, sl.l8[i - 6], 64, sl.use_default[i]))
            return rc;
    }
    return 0;
}

void skip_hrd(Bits &bs) {  // E.1.2
    const uint32_t cpb_cnt = bs.ue() + 1;
    bs.u(4);
    bs.u(4);
    for (uint32_t i = 0; i < cpb_cnt && !bs.over; i++) {
        bs.ue();
        bs.ue();
        bs.u(1);
    }
    bs.u(5);
    bs.u(5);
    bs.u(5);
    bs.u(5);
}

int parse_sps(Bits &bs, Sps *table) {  // 7.3.2.1.1
    Sps s;
    s.profile_idc = (int)bs.u(8);
    const uint32_t cflags = bs.u(8);
    s.constraint3 = (cflags >> 4) & 1;
    s.level_idc = (int)bs.u(8);
    const uint32_t id = bs.ue();
    if (id > 31) return h264g_set_error(H264G_E_ARG, "seq_parameter_set_id %u", id);
    const int p = s.profile_idc;
    if (p == 100 || p == 110 || p == 122 || p == 244 || p == 44 || p == 83 || p == 86 || p == 118 || p == 128 ||
        p == 138 || p == 139 || p == 134 || p == 135) {
        s.chroma_format_idc = (int)bs.ue();
        if (s.chroma_format_idc != 1) return h264g_set_error(H264G_E_ARG, "only 4:2:0 is supported");
        const uint32_t bdl = bs.ue(), bdc = bs.ue();
        if (bs.u(1)) return h264g_set_error(H264G_E_ARG, "lossless coding (qpprime_y_zero_transform_bypass_flag) is not supported");
        if (bdl || bdc) return h264g_set_error(H264G_E_ARG, "only 8-bit samples are supported");
        if (bs.u(1))
            if (int rc = parse_scaling_lists(bs, s.sl, 8)) return rc;
    }
    if (s.chroma_format_idc != 1) return h264g_set_error(H264G_E_ARG, "only 4:2:0 is supported");
    s.log2_max_frame_num = (int)bs.ue() + 4;
    s.poc_type = (int)bs.ue();
    if (s.poc_type == 0) {
        s.log2_max_poc_lsb = (int)bs.ue() + 4;
    } else if (s.poc_type == 1) {
        s.delta_poc_always_zero = (int)bs.u(1);
        s.offset_non_ref = bs.se();
        s.offset_top_to_bottom = bs.se();
        const uint32_t n = bs.ue();
        if (n > 255) return h264g_set_error(H264G_E_ARG, "num_ref_frames_in_pic_order_cnt_cycle %u", n);
        s.n_cycle = (int)n;
        for (int i = 0; i < s.n_cycle; i++) {
            s.offset_ref[i] = bs.se();
            s.expected_delta_cycle += s.offset_ref[i];  // ExpectedDeltaPerPicOrderCntCycle (7-12)
        }
    } else if (s.poc_type != 2) {
        return h264g_set_error(H264G_E_ARG, "pic_order_cnt_type %d", s.poc_type);
    }
    s.max_num_ref_frames = (int)bs.ue();
    s.gaps = (int)bs.u(1);
    s.width_mbs = (int)bs.ue() + 1;
    s.height_map_units = (int)bs.ue() + 1;
    s.frame_mbs_only = (int)bs.u(1);
    if (!s.frame_mbs_only) return h264g_set_error(H264G_E_ARG, "field coding (frame_mbs_only_flag 0) is not supported");
    s.direct_8x8_inference = (int)bs.u(1);
    if (bs.u(1)) {  // frame cropping: the planes stay whole, as in the reference
        bs.ue();
        bs.ue();
        bs.ue();
        bs.ue();
    }
    if (bs.u(1)) {  // vui_parameters (E.1.1), read for max_num_reorder_frames
        if (bs.u(1)) {
            if (bs.u(8) == 255) {
                bs.u(16);
                bs.u(16);
            }
        }
        if (bs.u(1)) bs.u(1);
        if (bs.u(1)) {
            bs.u(3);
            bs.u(1);
            if (bs.u(1)) {
                bs.u(8);
                bs.u(8);
                bs.u(8);
            }
        }
        if (bs.u(1)) {
            bs.ue();
            bs.ue();
        }
        if (bs.u(1)) {
            bs.u(32);
            bs.u(32);
            bs.u(1);
        }
        const uint32_t nal_hrd = bs.u(1);
        if (nal_hrd) skip_hrd(bs);
        const uint32_t vcl_hrd = bs.u(1);
        if (vcl_hrd) skip_hrd(bs);
        if (nal_hrd || vcl_hrd) bs.u(1);
        bs.u(1);  // pic_struct_present_flag
        if (bs.u(1)) {  // bitstream_restriction_flag
            bs.u(1);
            bs.ue();
            bs.ue();
            bs.ue();
            bs.ue();
            s.max_num_reorder_frames = (int)bs.ue();
            s.max_dec_frame_buffering = (int)bs.ue();
        }
    }
    if (bs.over) return h264g_set_error(H264G_E_ARG, "truncated SPS");
    s.valid = true;
    table[id] = s;
    return 0;
}

int parse_pps(Bits &bs, Pps *table, const Sps *sps) {  // 7.3.2.2
    Pps q;
    const uint32_t id = bs.ue();
    if (id > 255) return h264g_set_error(H264G_E_ARG, "pic_parameter_set_id %u", id);
    q.sps_id = (int)bs.ue();
    if (q.sps_id > 31 || !sps[q.sps_id].valid) return h264g_set_error(H264G_E_ARG, "PPS refers to a missing SPS");
    q.cabac = (int)bs.u(1);
    q.bottom_field_pic_order = (int)bs.u(1);
    q.num_slice_groups = (int)bs.ue() + 1;
    if (q.num_slice_groups != 1) return h264g_set_error(H264G_E_ARG, "slice groups are not supported");
    q.num_ref_idx_default[0] = (int)bs.ue() + 1;
    q.num_ref_idx_default[1] = (int)bs.ue() + 1;
    q.weighted_pred = (int)bs.u(1);
    q.weighted_bipred_idc = (int)bs.u(2);
    q.pic_init_qp = 26 + bs.se();
    bs.se();  // pic_init_qs
    q.chroma_qp_index_offset = bs.se();
    q.deblocking_filter_control = (int)bs.u(1);
    q.constrained_intra_pred = (int)bs.u(1);
    q.redundant_pic_cnt_present = (int)bs.u(1);
    q.second_chroma_qp_index_offset = q.chroma_qp_index_offset;
    if (bs.more_rbsp_data()) {
        q.transform_8x8_mode = (int)bs.u(1);
        // the reference reads second_chroma_qp_index_offset only inside the scaling-matrix
        // branch (H264PPS.cpp:207-227): without a matrix the PPS keeps chroma_qp_index_offset
        if (bs.u(1)) {
            if (int rc = parse_scaling_lists(bs, q.sl, 6 + 2 * q.transform_8x8_mode)) return rc;
            q.second_chroma_qp_index_offset = bs.se();
        }
    }
    if (bs.over) return h264g_set_error(H264G_E_ARG, "truncated PPS");
    q.valid = true;
    table[id] = q;
    return 0;
}

// ------------------------------------------------------------------ slice header
struct Mmco {
    int op, diff, lt_pic_num, lt_idx, max_lt_idx_plus1;
};
struct Slice {
    int nal_ref_idc = 0, idr = 0;
    int first_mb = 0, slice_type = 0, pps_id = 0, frame_num = 0, idr_pic_id = 0, poc_lsb = 0, delta_poc_bottom = 0;
    int delta_poc[2] = {0, 0};  // delta_pic_order_cnt[0..1] (pic_order_cnt_type 1)
    int redundant_pic_cnt = 0, direct_spatial = 1, num_ref_idx[2] = {0, 0};
    std::vector<std::pair<int, int>> mods[2];  // (modification_of_pic_nums_idc, value)
    int luma_log2_wd = 0, chroma_log2_wd = 0, luma_flag_last[2] = {0, 0};
    int16_t lw[2][32], lo[2][32], cw[2][32][2], co[2][32][2];
    int no_output_of_prior_pics = 0, long_term_reference = 0, adaptive_marking = 0;
    std::vector<Mmco> mmco;
    int cabac_init_idc = 0, slice_qp = 26, disable_dbk = 0, alpha_div2 = 0, beta_div2 = 0;
};

int parse_pred_weight_table(Bits &bs, Slice &s) {  // 7.3.3.2
    s.luma_log2_wd = (int)bs.ue();
    s.chroma_log2_wd = (int)bs.ue();
    for (int l = 0; l < (s.slice_type == H264G_SLICE_B ? 2 : 1); l++) {
        for (int i = 0; i < s.num_ref_idx[l]; i++) {
            const int lf = (int)bs.u(1);
            s.luma_flag_last[l] = lf;
            s.lw[l][i] = (int16_t)(1 << s.luma_log2_wd);
            s.lo[l][i] = 0;
            if (lf) {
                s.lw[l][i] = (int16_t)bs.se();
                s.lo[l][i] = (int16_t)bs.se();
            }
            const int cf = (int)bs.u(1);
            for (int c = 0; c < 2; c++) {
                s.cw[l][i][c] = (int16_t)(1 << s.chroma_log2_wd);
                s.co[l][i][c] = 0;
                if (cf) {
                    s.cw[l][i][c] = (int16_t)bs.se();
                    s.co[l][i][c] = (int16_t)bs.se();
                }
            }
        }
    }
    return 0;
}

int parse_slice_header(Bits &bs, Slice &s, int nal_type, int nal_ref_idc, const Sps *sps_t, const Pps *pps_t,
                       const Sps **sps_out, const Pps **pps_out) {  // 7.3.3
    s.nal_ref_idc = nal_ref_idc;
    s.idr = nal_type == 5;
    s.first_mb = (int)bs.ue();
    const uint32_t st = bs.ue() % 5;
    if (st == 3 || st == 4) return h264g_set_error(H264G_E_ARG, "SP / SI slices are not supported");
    s.slice_type = st == 0 ? H264G_SLICE_P : st == 1 ? H264G_SLICE_B : H264G_SLICE_I;
    s.pps_id = (int)bs.ue();
    if (s.pps_id > 255 || !pps_t[s.pps_id].valid) return h264g_set_error(H264G_E_ARG, "slice refers to a missing PPS");
    const Pps &pps = pps_t[s.pps_id];
    const Sps &sps = sps_t[pps.sps_id];
    *sps_out = &sps;
    *pps_out = &pps;
    s.frame_num = (int)bs.u(sps.log2_max_frame_num);
    if (s.idr) s.idr_pic_id = (int)bs.ue();
    if (sps.poc_type == 1 && !sps.delta_poc_always_zero) {
        s.delta_poc[0] = bs.se();
        if (pps.bottom_field_pic_order) s.delta_poc[1] = bs.se();
    }
    if (sps.poc_type == 0) {
        s.poc_lsb = (int)bs.u(sps.log2_max_poc_lsb);
        if (pps.bottom_field_pic_order) s.delta_poc_bottom = bs.se();
    }
    if (pps.redundant_pic_cnt_present) s.redundant_pic_cnt = (int)bs.ue();
    if (s.slice_type == H264G_SLICE_B) s.direct_spatial = (int)bs.u(1);
    s.num_ref_idx[0] = s.slice_type == H264G_SLICE_I ? 0 : pps.num_ref_idx_default[0];
    s.num_ref_idx[1] = s.slice_type == H264G_SLICE_B ? pps.num_ref_idx_default[1] : 0;
    if (s.slice_type != H264G_SLICE_I && bs.u(1)) {
        s.num_ref_idx[0] = (int)bs.ue() + 1;
        if (s.slice_type == H264G_SLICE_B) s.num_ref_idx[1] = (int)bs.ue() + 1;
    }
    if (s.num_ref_idx[0] > 32 || s.num_ref_idx[1] > 32) return h264g_set_error(H264G_E_ARG, "num_ref_idx_active > 32");
    for (int l = 0; l < (s.slice_type == H264G_SLICE_B ? 2 : s.slice_type == H264G_SLICE_P ? 1 : 0); l++) {
        if (bs.u(1)) {  // ref_pic_list_modification_flag_lX (7.3.3.1)
            for (int guard = 0; guard < 100; guard++) {
                const int idc = (int)bs.ue();
                if (idc == 3 || bs.over) break;
                if (idc > 2) return h264g_set_error(H264G_E_ARG, "modification_of_pic_nums_idc %d", idc);
                s.mods[l].push_back({idc, (int)bs.ue()});
            }
        }
    }
    if ((pps.weighted_pred && s.slice_type == H264G_SLICE_P) || (pps.weighted_bipred_idc == 1 && s.slice_type == H264G_SLICE_B))
        parse_pred_weight_table(bs, s);
    if (nal_ref_idc) {  // dec_ref_pic_marking (7.3.3.3)
        if (s.idr) {
            s.no_output_of_prior_pics = (int)bs.u(1);
            s.long_term_reference = (int)bs.u(1);
        } else if ((s.adaptive_marking = (int)bs.u(1))) {
            for (int guard = 0; guard < 100; guard++) {
                Mmco m{(int)bs.ue(), 0, 0, 0, 0};
                if (m.op == 0 || bs.over) break;
                if (m.op == 1 || m.op == 3) m.diff = (int)bs.ue() + 1;
                if (m.op == 2) m.lt_pic_num = (int)bs.ue();
                if (m.op == 3 || m.op == 6) m.lt_idx = (int)bs.ue();
                if (m.op == 4) m.max_lt_idx_plus1 = (int)bs.ue();
                if (m.op > 6) return h264g_set_error(H264G_E_ARG, "memory_management_control_operation %d", m.op);
                s.mmco.push_back(m);
            }
        }
    }
    if (pps.cabac && s.slice_type != H264G_SLICE_I) {
        s.cabac_init_idc = (int)bs.ue();
        if (s.cabac_init_idc > 2) return h264g_set_error(H264G_E_ARG, "cabac_init_idc %d", s.cabac_init_idc);
    }
    s.slice_qp = pps.pic_init_qp + bs.se();
    if (s.slice_qp < 0 || s.slice_qp > 51) return h264g_set_error(H264G_E_ARG, "SliceQPY %d", s.slice_qp);
    if (pps.deblocking_filter_control) {
        s.disable_dbk = (int)bs.ue();
        if (s.disable_dbk != 1) {
            s.alpha_div2 = bs.se();
            s.beta_div2 = bs.se();
        }
    }
    if (bs.over) return h264g_set_error(H264G_E_ARG, "truncated slice header");
    return 0;
}

// ------------------------------------------------------------------ macroblock layer
// Per-MB state the parse of later MBs reads: the TotalCoeff of each 4x4 block (luma in
// raster order of the MB's 4x4 grid, chroma [c][raster of the 2x2 grid]) for nC.
struct MbNz {
    uint8_t avail = 0;  // parsed in this picture
    uint8_t skip = 0, pcm = 0;
    uint8_t luma[16];
    uint8_t chroma[2][4];
};

// luma4x4BlkIdx -> raster index of the MB's 4x4 grid
inline int blk_raster(int b) { return (((b >> 3) & 1) * 2 + ((b >> 1) & 1)) * 4 + ((b >> 2) & 1) * 2 + (b & 1); }

struct Residual {
    // parsed coefficient blocks in scan order
    int16_t dc[16];            // Intra16x16DCLevel
    int16_t luma[16][16];      // by luma4x4BlkIdx (AC blocks: index 0..14 = scan 1..15)
    int16_t luma8[4][64];      // 8x8 blocks (CAVLC: interleaved 4x4 parses)
    int16_t cdc[2][4];
    int16_t cac[2][4][15];
};

// What deriving a parsed MB's motion vectors and intra modes reads of its syntax, kept
// when that derivation runs after the picture's slice data has parsed.
struct FinMb {
    uint8_t has_mvd;
    int16_t mvd[2][16][2];
    uint8_t isyn[16][2];
};

// One picture's slices as the stream sends them (header parsed, RBSP kept), the parameter
// sets in force for it, and its records.  Its slice data may be parsed ahead, on a thread
// of its own, into records whose motion vectors and modes wait for the in-order pass.
struct PicJob {
    struct SliceNal {
        Slice sl;
        std::vector<uint8_t> rbsp;
        size_t data_bit = 0;  // first bit of slice_data()
        bool over = false;    // the header read past the end of the RBSP
        int mb_end = 0;       // MBs covered once parsed ahead
        int damaged_mb = -1;  // the MB where its slice data failed, parsed ahead
        std::string damage;   // and why
    };
    std::vector<SliceNal> slices;
    Sps sps;
    Pps pps;
    int32_t rc = 0;           // an error met gathering the picture: returned in its place
    std::string err;
    std::vector<h264g_mb> mbs;
    std::vector<int16_t> levels;
    std::vector<FinMb> fin;
    bool ahead_ok = false;    // the slice data parsed ahead without an error
    bool derived = false;     // and its vectors and modes too (a picture without B slices)
    h264g_mvfield *field = nullptr;
    std::thread worker;
    ~PicJob() {
        if (worker.joinable()) worker.join();
        if (field) h264g_mvfield_destroy(field);
    }
};

}  // namespace

struct h264g_parser {
    std::vector<uint8_t> data;  // the byte stream (a copy)
    size_t pos = 0;             // next unread byte
    Sps sps[32];
    Pps pps[256];
    int width_mbs = 0, height_mbs = 0, dpb_slots = 0;
    // DPB: frames kept for reference
    struct Frame {
        int slot = -1, frame_num = 0, poc = 0, long_term_idx = -1;
        bool short_ref = false, long_ref = false;
        std::vector<h264g_mb> mbs;  // final records (co-located picture of direct prediction)
    };
    std::vector<Frame> dpb;
    int max_long_term_idx = -1;  // "no long-term frame indices"
    // POC type 0 / 2 state (8.2.1)
    int prev_poc_msb = 0, prev_poc_lsb = 0, prev_frame_num_offset = 0, prev_frame_num = 0;
    int prev_raw_frame_num = 0;  // frame_num of the previous picture as its slices sent it
    bool last_pic_mmco5 = false; // the previous picture (in decoding order) had MMCO 5
    bool prev_mmco5 = false;
    int prev_ref_top_poc = 0;
    h264g_mvfield *field = nullptr;
    int pictures = 0;              // pictures handed out
    // pictures gathered ahead of the one handed out (at most `threads`), the slice data of
    // each parsing on its own thread, and the picture handed out by the last
    // h264g_parser_next (its records stay valid until the next call)
    int threads = 1;
    int gathered = 0;
    bool gather_end = false;       // the stream (or an error) ended gathering
    std::deque<std::unique_ptr<PicJob>> ahead;
    std::unique_ptr<PicJob> out_job;
    char damage[512] = {0};     // why the last picture's damaged slice stopped ("" if none)
    ~h264g_parser() {
        if (field) h264g_mvfield_destroy(field);
    }
};

namespace {

// ------------------------------------------------------------------ CAVLC residual block
// residual_block_cavlc (7.3.5.3.2, 9.2): coeffLevel[startIdx..endIdx] of maxNumCoeff.
int residual_block(Bits &bs, int nC, int start, int end, int max_num, int16_t *out, int *total) {
    const Tables &T = tables();
    for (int i = 0; i < max_num; i++) out[i] = 0;
    int t1, tc;
    if (nC >= 8) {  // 6-bit fixed-length code: 0000 11 = (0, 0), else TotalCoeff - 1 (4 bits), TrailingOnes (2 bits)
        const uint32_t c = bs.u(6);
        t1 = c == 3 ? 0 : (int)(c & 3);
        tc = c == 3 ? 0 : (int)(c >> 2) + 1;
    } else {
        const Vlc::E *ct = T.coeff_token[nC == -1 ? 3 : nC < 2 ? 0 : nC < 4 ? 1 : 2].read(bs);
        if (!ct) return h264g_set_error(H264G_E_ARG, "invalid coeff_token");
        t1 = ct->a;
        tc = ct->b;
    }
    *total = tc;
    if (tc == 0) return 0;
    if (tc > max_num || t1 > tc) return h264g_set_error(H264G_E_ARG, "coeff_token beyond the block");
    int level[16];
    int suffix_length = (tc > 10 && t1 < 3) ? 1 : 0;
    for (int i = 0; i < tc; i++) {
        if (i < t1) {
            level[i] = bs.bit() ? -1 : 1;
            continue;
        }
        int prefix = 0;  // level_prefix (9.2.2.1)
        while (!bs.bit()) {
            if (++prefix > 31 || bs.over) return h264g_set_error(H264G_E_ARG, "invalid level_prefix");
        }
        int suffix_size = suffix_length;
        if (prefix == 14 && suffix_length == 0) suffix_size = 4;
        if (prefix >= 15) suffix_size = prefix - 3;
        int code = (std::min(15, prefix) << suffix_length) + (suffix_size ? (int)bs.u(suffix_size) : 0);
        if (prefix >= 15 && suffix_length == 0) code += 15;
        if (prefix >= 16) code += (1 << (prefix - 3)) - 4096;
        if (i == t1 && t1 < 3) code += 2;
        level[i] = (code & 1) ? (-code - 1) >> 1 : (code + 2) >> 1;
        if (suffix_length == 0) suffix_length = 1;
        if (std::abs(level[i]) > (3 << (suffix_length - 1)) && suffix_length < 6) suffix_length++;
    }
    int zeros_left = 0;
    if (tc < end - start + 1) {
        const Vlc::E *tz = max_num == 4 ? T.total_zeros_cdc[tc - 1].read(bs) : T.total_zeros[tc - 1].read(bs);
        if (!tz) return h264g_set_error(H264G_E_ARG, "invalid total_zeros");
        zeros_left = tz->a;
    }
    int run[16];
    for (int i = 0; i < tc - 1; i++) {
        run[i] = 0;
        if (zeros_left > 0) {
            const Vlc::E *rb = T.run_before[std::min(zeros_left, 7) - 1].read(bs);
            if (!rb)
                return h264g_set_error(H264G_E_ARG, "invalid run_before (nC %d, TotalCoeff %d, zerosLeft %d)", nC, tc,
                                       zeros_left);
            run[i] = rb->a;
            zeros_left -= run[i];
            if (zeros_left < 0) return h264g_set_error(H264G_E_ARG, "run_before beyond total_zeros");
        }
    }
    run[tc - 1] = zeros_left;
    int coeff = -1;
    for (int i = tc - 1; i >= 0; i--) {
        coeff += run[i] + 1;
        if (start + coeff > end) return h264g_set_error(H264G_E_ARG, "coefficients beyond the block");
        out[start + coeff] = (int16_t)level[i];
    }
    return bs.over ? h264g_set_error(H264G_E_ARG, "truncated residual") : 0;
}

// mb_type tables (7-11, 7-13, 7-14): partition shape and prediction of each partition
// (bit 0 = list 0, bit 1 = list 1) of P and B inter MB types
struct InterType { int shape; int pred[2]; };
const InterType kPType[5] = {{H264G_PART_16x16, {1, 0}}, {H264G_PART_16x8, {1, 1}}, {H264G_PART_8x16, {1, 1}},
                             {H264G_PART_8x8, {0, 0}}, {H264G_PART_8x8, {0, 0}}};
const InterType kBType[23] = {
    {H264G_PART_8x8, {0, 0}},   // B_Direct_16x16
    {H264G_PART_16x16, {1, 0}}, {H264G_PART_16x16, {2, 0}}, {H264G_PART_16x16, {3, 0}},
    {H264G_PART_16x8, {1, 1}},  {H264G_PART_8x16, {1, 1}},  {H264G_PART_16x8, {2, 2}}, {H264G_PART_8x16, {2, 2}},
    {H264G_PART_16x8, {1, 2}},  {H264G_PART_8x16, {1, 2}},  {H264G_PART_16x8, {2, 1}}, {H264G_PART_8x16, {2, 1}},
    {H264G_PART_16x8, {1, 3}},  {H264G_PART_8x16, {1, 3}},  {H264G_PART_16x8, {2, 3}}, {H264G_PART_8x16, {2, 3}},
    {H264G_PART_16x8, {3, 1}},  {H264G_PART_8x16, {3, 1}},  {H264G_PART_16x8, {3, 2}}, {H264G_PART_8x16, {3, 2}},
    {H264G_PART_16x8, {3, 3}},  {H264G_PART_8x16, {3, 3}},
    {H264G_PART_8x8, {0, 0}},   // B_8x8
};
// sub_mb_type (7-17, 7-18): shape and prediction; B index 0 is B_Direct_8x8
const int kPSubShape[4] = {H264G_SUB_8x8, H264G_SUB_8x4, H264G_SUB_4x8, H264G_SUB_4x4};
const int kBSubShape[13] = {H264G_SUB_8x8, H264G_SUB_8x8, H264G_SUB_8x8, H264G_SUB_8x8, H264G_SUB_8x4,
                            H264G_SUB_4x8, H264G_SUB_8x4, H264G_SUB_4x8, H264G_SUB_8x4, H264G_SUB_4x8,
                            H264G_SUB_4x4, H264G_SUB_4x4, H264G_SUB_4x4};
const int kBSubPred[13] = {0, 1, 2, 3, 1, 1, 2, 2, 3, 3, 1, 2, 3};

// 4x4 raster offsets (x, y in 4x4 units) of the partitions of a shape: quadrant q of
// 16x8 / 8x16 and the sub-partitions of a sub shape
inline int sub_parts(int shape) { return shape == H264G_SUB_8x8 ? 1 : shape == H264G_SUB_4x4 ? 4 : 2; }
inline void sub_origin(int shape, int k, int &x, int &y) {
    x = shape == H264G_SUB_4x8 ? k : shape == H264G_SUB_4x4 ? (k & 1) : 0;
    y = shape == H264G_SUB_8x4 ? k : shape == H264G_SUB_4x4 ? (k >> 1) : 0;
}

struct PicState {
    h264g_mvfield *field;          // motion field the MBs' vectors and modes go into, or
    std::vector<FinMb> *fin;       // (if non-null) where their syntax waits for it
    int picture;                   // the picture's index in the stream (messages)
    const Sps *sps;
    const Pps *pps;
    const Slice *sl;
    int W, H;
    std::vector<MbNz> nz;          // this slice's MBs (the others: not available)
    h264g_mb *mbs;                 // the picture's records
    std::vector<int16_t> *levels;
    uint32_t slice_num;            // the slice's index in its picture
    size_t mark = 0;               // levels->size() when the current MB began
    bool luma_done = false;        // the current MB's luma residual has parsed
    bool partial = false;          // the MB that failed holds a record of what it parsed
};

// TotalCoeff of the 4x4 luma block at (x, y) in 4x4 units of MB mb, or -1 if unavailable
// (outside the picture or not yet parsed); the neighbour MB is located from the
// current one (single-slice pictures: every parsed MB is available)
int nz_luma(const PicState &st, int mb, int bx, int by) {
    int mx = mb % st.W, my = mb / st.W;
    if (bx < 0) { bx += 4; mx--; }
    if (by < 0) { by += 4; my--; }
    if (mx < 0 || my < 0) return -1;
    const MbNz &n = st.nz[my * st.W + mx];
    if (!n.avail) return -1;
    if (n.skip) return 0;
    if (n.pcm) return 16;
    return n.luma[by * 4 + bx];
}
int nz_chroma(const PicState &st, int mb, int c, int bx, int by) {
    int mx = mb % st.W, my = mb / st.W;
    if (bx < 0) { bx += 2; mx--; }
    if (by < 0) { by += 2; my--; }
    if (mx < 0 || my < 0) return -1;
    const MbNz &n = st.nz[my * st.W + mx];
    if (!n.avail) return -1;
    if (n.skip) return 0;
    if (n.pcm) return 16;
    return n.chroma[c][by * 2 + bx];
}
inline int nc_of(int a, int b) {  // 9.2.1
    if (a >= 0 && b >= 0) return (a + b + 1) >> 1;
    if (a >= 0) return a;
    if (b >= 0) return b;
    return 0;
}

// Appends the masks and values of the MB's blocks in the order h264g.h documents.
void emit_levels(const Residual &r, bool i16, bool t8, int cbp, std::vector<int16_t> &lv) {
    uint16_t masks[27];
    int16_t vals[27 * 16];
    int nm = 0, nv = 0;
    auto block = [&](const int16_t *c, int n, int first_pos) {  // n coefficients at scan first_pos..
        uint32_t m = 0;
        for (int i = 0; i < n; i++) m |= (uint32_t)(c[i] != 0) << i;  // branch-free, vectorisable
        for (uint32_t t = m; t; t &= t - 1) vals[nv++] = c[__builtin_ctz(t)];
        masks[nm++] = (uint16_t)(m << first_pos);
    };
    if (i16) block(r.dc, 16, 0);
    for (int q = 0; q < 4; q++) {
        if (!((cbp >> q) & 1)) continue;
        if (t8) {
            for (int k = 0; k < 4; k++) block(&r.luma8[q][16 * k], 16, 0);
        } else {
            for (int k = 0; k < 4; k++) {
                if (i16) block(r.luma[4 * q + k], 15, 1);
                else block(r.luma[4 * q + k], 16, 0);
            }
        }
    }
    if ((cbp >> 4) >= 1)
        for (int c = 0; c < 2; c++) block(r.cdc[c], 4, 0);
    if ((cbp >> 4) == 2)
        for (int c = 0; c < 2; c++)
            for (int b = 0; b < 4; b++) block(r.cac[c][b], 15, 1);
    for (int i = 0; i < nm; i++) lv.push_back((int16_t)masks[i]);
    lv.insert(lv.end(), vals, vals + nv);
}

// residual() (7.3.5.3) for CAVLC, 4:2:0
int parse_residual(Bits &bs, PicState &st, int mb, bool i16, bool t8, int cbp, Residual &r, MbNz &nzc) {
    memset(&r, 0, sizeof(r));
    memset(nzc.luma, 0, sizeof(nzc.luma));
    memset(nzc.chroma, 0, sizeof(nzc.chroma));
    int total = 0;
    auto luma_nc = [&](int blk) {
        const int rr = blk_raster(blk), bx = rr & 3, by = rr >> 2;
        const int a = bx > 0 ? nzc.luma[by * 4 + bx - 1] : nz_luma(st, mb, -1, by);
        const int b = by > 0 ? nzc.luma[(by - 1) * 4 + bx] : nz_luma(st, mb, bx, -1);
        return nc_of(a, b);
    };
    if (i16) {
        if (int rc = residual_block(bs, luma_nc(0), 0, 15, 16, r.dc, &total)) return rc;
    }
    for (int q = 0; q < 4; q++) {
        for (int k = 0; k < 4; k++) {
            const int blk = 4 * q + k;
            if (!((cbp >> q) & 1)) continue;
            int16_t c[16];
            const int nc = luma_nc(blk);
            if (i16) {
                if (int rc = residual_block(bs, nc, 0, 14, 15, c, &total)) return rc;
                memcpy(r.luma[blk], c, sizeof(c));
            } else {
                if (int rc = residual_block(bs, nc, 0, 15, 16, c, &total)) return rc;
                memcpy(r.luma[blk], c, sizeof(c));
                if (t8)
                    for (int i = 0; i < 16; i++) r.luma8[q][4 * i + k] = c[i];
            }
            nzc.luma[blk_raster(blk)] = (uint8_t)total;
        }
    }
    st.luma_done = true;
    const int cc = cbp >> 4;
    for (int c = 0; c < 2 && cc; c++) {
        if (int rc = residual_block(bs, -1, 0, 3, 4, r.cdc[c], &total)) return rc;
    }
    for (int c = 0; c < 2; c++)
        for (int b = 0; b < 4; b++) {
            if (cc != 2) continue;
            const int bx = b & 1, by = b >> 1;
            const int a = bx > 0 ? nzc.chroma[c][by * 2 + bx - 1] : nz_chroma(st, mb, c, -1, by);
            const int bb = by > 0 ? nzc.chroma[c][(by - 1) * 2 + bx] : nz_chroma(st, mb, c, bx, -1);
            int16_t co[16];
            if (int rc = residual_block(bs, nc_of(a, bb), 0, 14, 15, co, &total)) return rc;
            memcpy(r.cac[c][b], co, 15 * sizeof(int16_t));
            nzc.chroma[c][b] = (uint8_t)total;
        }
    return 0;
}

// One macroblock_layer (7.3.5) of a P / B / I slice into m (mvd and intra syntax beside it).
int parse_mb(Bits &bs, PicState &st, int mb, int &qp, h264g_mb &m, int16_t (&mvd)[2][16][2], uint8_t (&isyn)[16][2],
             bool &has_mvd) {
    const Slice &sl = *st.sl;
    const Pps &pps = *st.pps;
    MbNz &nzc = st.nz[mb];
    memset(&m, 0, sizeof(m));
    memset(mvd, 0, sizeof(mvd));
    has_mvd = false;
    for (int l = 0; l < 2; l++)
        for (int q = 0; q < 4; q++) m.ref_idx[l][q] = -1;
    m.level_offset = (uint32_t)st.levels->size();
    int mb_type = (int)bs.ue();
    bool intra = sl.slice_type == H264G_SLICE_I;
    if (sl.slice_type == H264G_SLICE_P) {
        if (mb_type >= 5) { intra = true; mb_type -= 5; }
    } else if (sl.slice_type == H264G_SLICE_B) {
        if (mb_type >= 23) { intra = true; mb_type -= 23; }
    }
    if (intra && mb_type > 25) return h264g_set_error(H264G_E_ARG, "mb_type %d", mb_type);
    if (!intra && sl.slice_type == H264G_SLICE_P && mb_type > 4) return h264g_set_error(H264G_E_ARG, "mb_type %d", mb_type);
    nzc.avail = 1;
    nzc.skip = 0;
    nzc.pcm = 0;
    if (intra && mb_type == 25) {  // I_PCM
        while (!bs.aligned()) bs.bit();
        m.mb_type = H264G_MB_IPCM;
        m.qp_y = (uint8_t)qp;
        for (int i = 0; i < 384; i++) st.levels->push_back((int16_t)bs.u(8));
        nzc.pcm = 1;
        return bs.over ? h264g_set_error(H264G_E_ARG, "truncated I_PCM") : 0;
    }
    bool t8 = false, i16 = false;
    int cbp = 0;
    bool no_sub_lt8 = true;
    const int nref[2] = {sl.num_ref_idx[0], sl.num_ref_idx[1]};
    if (intra) {
        if (mb_type == 0) {  // I_NxN
            if (pps.transform_8x8_mode) t8 = bs.u(1);
            const int n = t8 ? 4 : 16;
            for (int b = 0; b < n; b++) {
                isyn[b][0] = (uint8_t)bs.u(1);
                isyn[b][1] = isyn[b][0] ? 0 : (uint8_t)bs.u(3);
            }
            m.mb_type = t8 ? H264G_MB_I8x8 : H264G_MB_I4x4;
        } else {
            i16 = true;
            m.mb_type = H264G_MB_I16x16;
            m.intra16x16_pred_mode = (uint8_t)((mb_type - 1) % 4);
            cbp = (((mb_type - 1) / 4) % 3) << 4 | (mb_type >= 13 ? 15 : 0);
        }
        m.intra_chroma_pred_mode = (uint8_t)bs.ue();
        if (m.intra_chroma_pred_mode > 3) return h264g_set_error(H264G_E_ARG, "intra_chroma_pred_mode");
    } else if (sl.slice_type == H264G_SLICE_P || mb_type != 0) {
        const bool is_b = sl.slice_type == H264G_SLICE_B;
        const InterType &it = is_b ? kBType[mb_type] : kPType[mb_type];
        m.mb_type = is_b ? H264G_MB_B : H264G_MB_P;
        m.part_shape = (uint8_t)it.shape;
        has_mvd = true;
        if (it.shape == H264G_PART_8x8) {  // sub_mb_pred (7.3.5.2)
            int sub[4];
            for (int q = 0; q < 4; q++) {
                sub[q] = (int)bs.ue();
                if (sub[q] > (is_b ? 12 : 3)) return h264g_set_error(H264G_E_ARG, "sub_mb_type %d", sub[q]);
                m.sub_part_shape[q] = (uint8_t)(is_b ? kBSubShape[sub[q]] : kPSubShape[sub[q]]);
                if (is_b && sub[q] == 0) {
                    m.flags |= H264G_MBF_DIRECT(q);
                    if (!st.sps->direct_8x8_inference) no_sub_lt8 = false;
                } else if (sub_parts(m.sub_part_shape[q]) > 1) {
                    no_sub_lt8 = false;
                }
            }
            auto pred = [&](int q) { return is_b ? kBSubPred[sub[q]] : 1; };
            for (int l = 0; l < 2; l++)
                for (int q = 0; q < 4; q++) {
                    if (!(pred(q) & (1 << l))) continue;
                    m.ref_idx[l][q] = 0;
                    if (nref[l] > 1 && !(mb_type == 4 && !is_b)) m.ref_idx[l][q] = (int8_t)bs.te(nref[l] - 1);
                    if (m.ref_idx[l][q] >= nref[l]) return h264g_set_error(H264G_E_ARG, "ref_idx out of range");
                }
            for (int l = 0; l < 2; l++)
                for (int q = 0; q < 4; q++) {
                    if (!(pred(q) & (1 << l))) continue;
                    const int shape = m.sub_part_shape[q];
                    for (int k = 0; k < sub_parts(shape); k++) {
                        int x, y;
                        sub_origin(shape, k, x, y);
                        const int r = ((q >> 1) * 2 + y) * 4 + (q & 1) * 2 + x;
                        mvd[l][r][0] = (int16_t)bs.se();
                        mvd[l][r][1] = (int16_t)bs.se();
                    }
                }
        } else {  // mb_pred of an inter MB (7.3.5.1)
            const int np = it.shape == H264G_PART_16x16 ? 1 : 2;
            auto quads = [&](int k, int *q) -> int {  // quadrants of partition k
                if (np == 1) { q[0] = 0; q[1] = 1; q[2] = 2; q[3] = 3; return 4; }
                if (it.shape == H264G_PART_16x8) { q[0] = 2 * k; q[1] = 2 * k + 1; return 2; }
                q[0] = k; q[1] = k + 2;
                return 2;
            };
            for (int l = 0; l < 2; l++)
                for (int k = 0; k < np; k++) {
                    if (!(it.pred[k] & (1 << l))) continue;
                    int ri = 0;
                    if (nref[l] > 1) ri = bs.te(nref[l] - 1);
                    if (ri >= nref[l]) return h264g_set_error(H264G_E_ARG, "ref_idx out of range");
                    int q[4];
                    const int nq = quads(k, q);
                    for (int j = 0; j < nq; j++) m.ref_idx[l][q[j]] = (int8_t)ri;
                }
            for (int l = 0; l < 2; l++)
                for (int k = 0; k < np; k++) {
                    if (!(it.pred[k] & (1 << l))) continue;
                    const int r = np == 1 ? 0 : it.shape == H264G_PART_16x8 ? 8 * k : 2 * k;
                    mvd[l][r][0] = (int16_t)bs.se();
                    mvd[l][r][1] = (int16_t)bs.se();
                }
        }
    } else {  // B_Direct_16x16
        m.mb_type = H264G_MB_B;
        m.part_shape = H264G_PART_8x8;
        m.flags |= H264G_MBF_DIRECT_ALL;
        if (!st.sps->direct_8x8_inference) no_sub_lt8 = false;
    }
    if (!i16) {
        const uint32_t code = bs.ue();
        if (code > 47) return h264g_set_error(H264G_E_ARG, "coded_block_pattern");
        const bool nxn = m.mb_type == H264G_MB_I4x4 || m.mb_type == H264G_MB_I8x8;
        cbp = nxn ? kCbpIntra[code] : kCbpInter[code];  // CodedBlockPatternLuma | Chroma << 4
        if ((cbp & 15) && pps.transform_8x8_mode && !nxn && no_sub_lt8 && !intra) t8 = bs.u(1);
    }
    if (t8) m.flags |= H264G_MBF_T8x8;
    m.cbp = (uint8_t)cbp;
    if ((cbp & 15) || (cbp >> 4) || i16) {
        // out-of-range deltas are clipped, as the reference does (H264MacroBlock.cpp:1116-1119)
        const int dq = std::min(25, std::max(-26, bs.se()));
        qp = (qp + dq + 52) % 52;  // 7-37
        Residual r;
        st.luma_done = false;
        if (int rc = parse_residual(bs, st, mb, i16, t8, cbp, r, nzc)) {
            // The reference reconstructs an MB whose residual fails to parse from the syntax
            // it holds (H264SliceData.cpp:375-379 only logs the failure): QPY keeps its reset
            // value 0, as it is set after residual() (H264MacroBlock.cpp:1110-1124); luma
            // levels are kept only once every luma block parsed (H264MacroBlock.cpp:1616-1622),
            // chroma blocks one by one as they parsed (H264MacroBlock.cpp:1626-1700).
            if (!st.luma_done) {
                memset(r.dc, 0, sizeof(r.dc));
                memset(r.luma, 0, sizeof(r.luma));
                memset(r.luma8, 0, sizeof(r.luma8));
            }
            m.qp_y = 0;
            emit_levels(r, i16, t8, cbp, *st.levels);
            st.partial = true;
            return rc;
        }
        emit_levels(r, i16, t8, cbp, *st.levels);
    } else {
        memset(nzc.luma, 0, sizeof(nzc.luma));
        memset(nzc.chroma, 0, sizeof(nzc.chroma));
    }
    m.qp_y = (uint8_t)qp;
    return bs.over ? h264g_set_error(H264G_E_ARG, "truncated macroblock") : 0;
}

// A skipped MB's record (P_Skip, or B_Skip: direct prediction of all four quadrants).
void skip_record(h264g_mb &m, int slice_type, int qp, uint32_t level_offset) {
    memset(&m, 0, sizeof(m));
    for (int l = 0; l < 2; l++)
        for (int q = 0; q < 4; q++) m.ref_idx[l][q] = -1;
    m.qp_y = (uint8_t)qp;
    m.level_offset = level_offset;
    if (slice_type == H264G_SLICE_P) {
        m.mb_type = H264G_MB_P_SKIP;
        m.ref_idx[0][0] = m.ref_idx[0][1] = m.ref_idx[0][2] = m.ref_idx[0][3] = 0;
    } else {
        m.mb_type = H264G_MB_B;
        m.part_shape = H264G_PART_8x8;
        m.flags = H264G_MBF_SKIP | H264G_MBF_DIRECT_ALL;
    }
}

// Intra prediction modes and motion vectors of a parsed MB (the motion field, in MB order).
int derive_mb(h264g_mvfield *field, int cip, int mb, h264g_mb &m, const int16_t *mvd, const uint8_t *isyn) {
    if (m.mb_type == H264G_MB_I4x4 || m.mb_type == H264G_MB_I8x8) {
        if (int rc = h264g_mvfield_intra_modes(field, (uint32_t)mb, &m, isyn, cip)) return rc;
    }
    const bool all_direct = (m.flags & H264G_MBF_DIRECT_ALL) == H264G_MBF_DIRECT_ALL;
    return h264g_mvfield_derive(field, (uint32_t)mb, &m, all_direct ? nullptr : mvd);
}

// A parsed MB joins its slice; its modes and vectors are derived now, or its syntax is
// kept for the pass that derives them in MB order.
int finish_mb(PicState &st, int mb, h264g_mb &m, const int16_t *mvd, const uint8_t *isyn) {
    m.slice_num = st.slice_num;
    if (st.fin) {
        FinMb &f = (*st.fin)[mb];
        f.has_mvd = mvd != nullptr;
        if (mvd) memcpy(f.mvd, mvd, sizeof(f.mvd));
        if (isyn && (m.mb_type == H264G_MB_I4x4 || m.mb_type == H264G_MB_I8x8)) memcpy(f.isyn, isyn, sizeof(f.isyn));
        return 0;
    }
    return derive_mb(st.field, st.pps->constrained_intra_pred, mb, m, mvd, isyn);
}

// An MB left undecoded (H264G_MBF_LOST): zero samples as I_PCM with QPY 0, the reset
// state the reference leaves such MBs in (H264PictureBase.cpp:45-67).  An MB where slice
// data failed stays in its slice; MBs no slice reached belong to no slice
// (H264G_SLICE_NONE).
void undecoded_record(h264g_mb *mbs, std::vector<int16_t> &levels, int mb, uint32_t slice_num) {
    h264g_mb &m = mbs[mb];
    memset(&m, 0, sizeof(m));
    for (int l = 0; l < 2; l++)
        for (int q = 0; q < 4; q++) m.ref_idx[l][q] = -1;
    m.mb_type = H264G_MB_IPCM;
    m.flags = H264G_MBF_LOST;
    m.slice_num = slice_num;
    m.level_offset = (uint32_t)levels.size();
    levels.resize(levels.size() + 384, 0);
}

int undecoded_mb(h264g_mb *mbs, std::vector<int16_t> &levels, h264g_mvfield *field, int mb, uint32_t slice_num) {
    undecoded_record(mbs, levels, mb, slice_num);
    return h264g_mvfield_derive(field, (uint32_t)mb, &mbs[mb], nullptr);
}

int mb_error(const PicState &st, int mb, int rc) {
    char msg[512];
    snprintf(msg, sizeof(msg), "%s", h264g_last_error());
    return h264g_set_error(rc, "picture %d MB %d: %s", st.picture, mb, msg);
}

// slice_data() (7.3.4) with CAVLC: mb_skip_run, then macroblock_layer; mb = MBs covered.
int parse_slice_data_cavlc(Bits &bs, PicState &st, int &mb) {
    const Slice &sl = *st.sl;
    const int n_mbs = st.W * st.H;
    int qp = sl.slice_qp;
    bool more = true;
    int16_t mvd[2][16][2];
    uint8_t isyn[16][2];
    while (more && mb < n_mbs) {
        st.mark = st.levels->size();
        if (sl.slice_type != H264G_SLICE_I) {
            const uint32_t run = bs.ue();
            if (run > (uint32_t)(n_mbs - mb)) return h264g_set_error(H264G_E_ARG, "mb_skip_run past the picture");
            for (uint32_t k = 0; k < run; k++, mb++) {
                h264g_mb &m = st.mbs[mb];
                skip_record(m, sl.slice_type, qp, (uint32_t)st.levels->size());
                st.nz[mb].avail = 1;
                st.nz[mb].skip = 1;
                if (int rc = finish_mb(st, mb, m, nullptr, nullptr)) return mb_error(st, mb, rc);
            }
            if (run > 0) more = bs.more_rbsp_data();
            if (!more || mb >= n_mbs) break;
        }
        h264g_mb &m = st.mbs[mb];
        bool has_mvd = false;
        if (int rc = parse_mb(bs, st, mb, qp, m, mvd, isyn, has_mvd)) {
            const int err = mb_error(st, mb, rc);
            if (st.partial) {  // reconstructed from what parsed; the slice ends with it
                m.flags |= H264G_MBF_LOST;
                if (finish_mb(st, mb, m, has_mvd ? &mvd[0][0][0] : nullptr, &isyn[0][0])) st.partial = false;
            }
            return err;
        }
        if (int rc = finish_mb(st, mb, m, has_mvd ? &mvd[0][0][0] : nullptr, &isyn[0][0])) return mb_error(st, mb, rc);
        mb++;
        more = bs.more_rbsp_data();
    }
    return 0;
}

// ------------------------------------------------------------------ CABAC (9.3)
#include "cabac_init.inc"

// Table 9-44 and the state transitions of 9.3.3.2.1.1
const uint8_t kRangeLps[64][4] = {
    {128, 176, 208, 240}, {128, 167, 197, 227}, {128, 158, 187, 216}, {123, 150, 178, 205}, {116, 142, 169, 195},
    {111, 135, 160, 185}, {105, 128, 152, 175}, {100, 122, 144, 166}, {95, 116, 137, 158}, {90, 110, 130, 150},
    {85, 104, 123, 142},  {81, 99, 117, 135},   {77, 94, 111, 128},   {73, 89, 105, 122},   {69, 85, 100, 116},
    {66, 80, 95, 110},    {62, 76, 90, 104},    {59, 72, 86, 99},     {56, 69, 81, 94},     {53, 65, 77, 89},
    {51, 62, 73, 85},     {48, 59, 69, 80},     {46, 56, 66, 76},     {43, 53, 63, 72},     {41, 50, 59, 69},
    {39, 48, 56, 65},     {37, 45, 54, 62},     {35, 43, 51, 59},     {33, 41, 48, 56},     {32, 39, 46, 53},
    {30, 37, 43, 50},     {29, 35, 41, 48},     {27, 33, 39, 45},     {26, 31, 37, 43},     {24, 30, 35, 41},
    {23, 28, 33, 39},     {22, 27, 32, 37},     {21, 26, 30, 35},     {20, 24, 29, 33},     {19, 23, 27, 31},
    {18, 22, 26, 30},     {17, 21, 25, 28},     {16, 20, 23, 27},     {15, 19, 22, 25},     {14, 18, 21, 24},
    {14, 17, 20, 23},     {13, 16, 19, 22},     {12, 15, 18, 21},     {12, 14, 17, 20},     {11, 14, 16, 19},
    {11, 13, 15, 18},     {10, 12, 15, 17},     {10, 12, 14, 16},     {9, 11, 13, 15},      {9, 11, 12, 14},
    {8, 10, 12, 14},      {8, 9, 11, 13},       {7, 9, 11, 12},       {7, 9, 10, 12},       {7, 8, 10, 11},
    {6, 8, 9, 11},        {6, 7, 9, 10},        {6, 7, 8, 9},         {2, 2, 2, 2}};
const uint8_t kTransLps[64] = {0,  0,  1,  2,  2,  4,  4,  5,  6,  7,  8,  9,  9,  11, 11, 12, 13, 13, 15, 15, 16, 16,
                               18, 18, 19, 19, 21, 21, 22, 22, 23, 24, 24, 25, 26, 26, 27, 27, 28, 29, 29, 30, 30, 30,
                               31, 32, 32, 33, 33, 33, 34, 34, 35, 35, 35, 36, 36, 36, 37, 37, 37, 38, 38, 63};
// ctxIdxInc of significant_coeff_flag / last_significant_coeff_flag of 8x8 blocks in frame
// MBs by scan position (Table 9-43)
const uint8_t kSig8x8[63] = {0,  1,  2,  3,  4,  5,  5,  4,  4,  3,  3,  4,  4,  4,  5,  5,  4,  4,  4,  4,  3,
                             3,  6,  7,  7,  7,  8,  9,  10, 9,  8,  7,  7,  6,  11, 12, 13, 11, 6,  7,  8,  9,
                             14, 10, 9,  8,  6,  11, 12, 13, 11, 6,  9,  14, 10, 9,  11, 12, 13, 11, 14, 10, 12};
const uint8_t kLast8x8[63] = {0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2,
                              3, 3, 3, 3, 3, 3, 3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 5, 5, 5, 5, 6, 6, 6, 6, 7, 7, 7, 7, 8, 8, 8};

// The arithmetic decoding engine (9.3.1.2, 9.3.3.2), one bit of the stream at a time.
struct Cabac {
    Bits *bs = nullptr;
    uint32_t range = 0, offset = 0;
    uint8_t state[460], mps[460];
    void init_contexts(int slice_type, int idc, int qp) {  // 9.3.1.1
        const int col = slice_type == H264G_SLICE_I ? 0 : 1 + idc;
        const int q = std::min(51, std::max(0, qp));
        for (int c = 0; c < 460; c++) {
            const int m = kCabacInit[c][col][0], n = kCabacInit[c][col][1];
            const int pre = std::min(126, std::max(1, ((m * q) >> 4) + n));
            state[c] = (uint8_t)(pre <= 63 ? 63 - pre : pre - 64);
            mps[c] = (uint8_t)(pre <= 63 ? 0 : 1);
        }
    }
    void init_engine() {
        range = 510;
        offset = bs->u(9);
    }
    int decision(int c) {  // 9.3.3.2.1
        const uint32_t lps = kRangeLps[state[c]][(range >> 6) & 3];
        range -= lps;
        int bin;
        if (offset >= range) {
            bin = !mps[c];
            offset -= range;
            range = lps;
            if (state[c] == 0) mps[c] = (uint8_t)(1 - mps[c]);
            state[c] = kTransLps[state[c]];
        } else {
            bin = mps[c];
            if (state[c] < 62) state[c]++;
        }
        renorm();
        return bin;
    }
    void renorm() {  // RenormD: shift range back to 9 bits, the same number of bits into offset
        if (range < 256) {
            const int n = __builtin_clz(range) - 23;
            range <<= n;
            offset = (offset << n) | bs->u(n);
        }
    }
    int bypass() {  // 9.3.3.2.3
        offset = (offset << 1) | bs->bit();
        if (offset >= range) {
            offset -= range;
            return 1;
        }
        return 0;
    }
    int terminate() {  // 9.3.3.2.2.3
        range -= 2;
        if (offset >= range) return 1;
        renorm();
        return 0;
    }
};

// What later MBs' context derivation reads of an MB (9.3.3.1.1).
struct CMb {
    uint8_t avail, skip, intra, pcm, i16, nxn, bdirect16, t8, cbp, chroma_mode, qp_delta_nz;
    uint8_t cbf_dc, cbf_cdc;   // Intra16x16 DC; chroma DC bit c
    uint16_t cbf_luma;         // 4x4 luma blocks, raster bit
    uint8_t cbf_cac[2];        // chroma AC, raster bit of the 2x2 grid
    int8_t ref[2][16];         // refIdx of each 4x4 block, -1 where the list is not used
    uint8_t direct[16];        // predicted by direct prediction (B_Skip / B_Direct / direct 8x8)
    uint8_t amvd[2][16][2];    // absolute mvd components (clipped to 255)
};

struct CabacPic {
    Cabac c;
    std::vector<CMb> mbs;
    int W = 0, H = 0;
    const CMb *mb_at(int mb, int dx, int dy) const {  // neighbour MB (dx, dy in {-1, 0}) or null
        const int x = mb % W + dx, y = mb / W + dy;
        if (x < 0 || y < 0 || x >= W) return nullptr;
        const CMb &n = mbs[y * W + x];
        return n.avail ? &n : nullptr;
    }
};

// The MB and raster 4x4 block holding 4x4 block (bx, by) of MB mb, bx / by in -1..3
// (the left / above neighbours); null when outside the picture or not yet decoded.
const CMb *blk_at(const CabacPic &cp, int mb, const CMb &cur, int bx, int by, int &r) {
    int dx = 0, dy = 0;
    if (bx < 0) { bx += 4; dx = -1; }
    if (by < 0) { by += 4; dy = -1; }
    r = by * 4 + bx;
    if (!dx && !dy) return &cur;
    return cp.mb_at(mb, dx, dy);
}

int cabac_mb_skip(CabacPic &cp, int mb, int slice_type) {
    const CMb *a = cp.mb_at(mb, -1, 0), *b = cp.mb_at(mb, 0, -1);
    const int inc = (a && !a->skip) + (b && !b->skip);
    return cp.c.decision((slice_type == H264G_SLICE_P ? 11 : 24) + inc);
}

// mb_type of I MBs, ctxIdxOffset 3 (I slices) or the intra suffix 17 / 32 (P / B slices);
// returns 0 (I_NxN), 1..24 (I_16x16) or 25 (I_PCM)
int cabac_mb_type_intra(CabacPic &cp, int mb, int base, bool islice) {
    Cabac &c = cp.c;
    if (islice) {
        const CMb *a = cp.mb_at(mb, -1, 0), *b = cp.mb_at(mb, 0, -1);
        const int inc = (a && !a->nxn) + (b && !b->nxn);
        if (!c.decision(base + inc)) return 0;
        base += 2;
    } else if (!c.decision(base)) {
        return 0;
    }
    if (c.terminate()) return 25;
    int t = 1 + 12 * c.decision(base + 1);
    if (c.decision(base + 2)) t += 4 + 4 * c.decision(base + 2 + (islice ? 1 : 0));
    t += 2 * c.decision(base + 3 + (islice ? 1 : 0));
    t += c.decision(base + 3 + (islice ? 2 : 0));
    return t;
}

int cabac_mb_type_p(CabacPic &cp, int mb, bool &intra) {
    Cabac &c = cp.c;
    intra = false;
    if (!c.decision(14)) {
        if (!c.decision(15)) return 3 * c.decision(16);  // P_L0_16x16 (0) / P_8x8 (3)
        return 2 - c.decision(17);                       // P_L0_L0_8x16 (2) / P_L0_L0_16x8 (1)
    }
    intra = true;
    return cabac_mb_type_intra(cp, mb, 17, false);
}

int cabac_mb_type_b(CabacPic &cp, int mb, bool &intra) {
    Cabac &c = cp.c;
    intra = false;
    const CMb *a = cp.mb_at(mb, -1, 0), *b = cp.mb_at(mb, 0, -1);
    const int inc = (a && !a->bdirect16) + (b && !b->bdirect16);
    if (!c.decision(27 + inc)) return 0;  // B_Direct_16x16
    if (!c.decision(27 + 3)) return 1 + c.decision(27 + 5);
    int bits = c.decision(27 + 4) << 3;
    bits |= c.decision(27 + 5) << 2;
    bits |= c.decision(27 + 5) << 1;
    bits |= c.decision(27 + 5);
    if (bits < 8) return bits + 3;
    if (bits == 13) {
        intra = true;
        return cabac_mb_type_intra(cp, mb, 32, false);
    }
    if (bits == 14) return 11;
    if (bits == 15) return 22;
    bits = (bits << 1) | c.decision(27 + 5);
    return bits - 4;
}

int cabac_sub_p(Cabac &c) {
    if (c.decision(21)) return 0;
    if (!c.decision(22)) return 1;
    return c.decision(23) ? 2 : 3;
}

int cabac_sub_b(Cabac &c) {
    if (!c.decision(36)) return 0;
    if (!c.decision(37)) return 1 + c.decision(39);
    int t = 3;
    if (c.decision(38)) {
        if (c.decision(39)) return 11 + c.decision(39);
        t += 4;
    }
    t += 2 * c.decision(39);
    t += c.decision(39);
    return t;
}

// ref_idx_lX of the partition whose top-left 4x4 block is (bx, by) (9.3.3.1.1.6)
int cabac_ref_idx(CabacPic &cp, int mb, const CMb &cur, int l, int bx, int by, bool bslice) {
    int inc = 0;
    for (int k = 0; k < 2; k++) {
        int r;
        const CMb *n = blk_at(cp, mb, cur, bx - (k == 0), by - (k == 1), r);
        if (!n || n->skip || n->intra) continue;
        if (bslice && n->direct[r]) continue;
        if (n->ref[l][r] > 0) inc += k ? 2 : 1;
    }
    int ref = 0;
    while (cp.c.decision(54 + inc)) {
        ref++;
        inc = ref == 1 ? 4 : 5;
        if (ref > 31) return -1;
    }
    return ref;
}

// one mvd component (comp 0: x, ctxIdxOffset 40; 1: y, 47) (9.3.3.1.1.7, UEG3 suffix)
int cabac_mvd(CabacPic &cp, int mb, const CMb &cur, int l, int bx, int by, int comp) {
    int sum = 0;
    for (int k = 0; k < 2; k++) {
        int r;
        const CMb *n = blk_at(cp, mb, cur, bx - (k == 0), by - (k == 1), r);
        if (n) sum += n->amvd[l][r][comp];
    }
    const int base = comp ? 47 : 40;
    Cabac &c = cp.c;
    if (!c.decision(base + (sum < 3 ? 0 : sum > 32 ? 2 : 1))) return 0;
    int v = 1, ctx = base + 3;
    while (v < 9 && c.decision(ctx)) {
        v++;
        if (ctx < base + 6) ctx++;
    }
    if (v >= 9) {
        int k = 3;
        while (c.bypass()) {
            v += 1 << k;
            if (++k > 24) return 0x7fffffff;
        }
        while (k--) v += c.bypass() << k;
    }
    return c.bypass() ? -v : v;
}

// coded_block_flag neighbour term (9.3.3.1.1.9); kind: 0 luma DC, 1 luma 4x4 (raster r),
// 2 chroma DC (component cc), 3 chroma AC (component cc, raster r of the 2x2 grid)
int cbf_term(const CMb *n, bool cur_intra, int kind, int r, int cc) {
    if (!n) return cur_intra ? 1 : 0;
    if (n->pcm) return 1;
    if (n->skip) return 0;
    switch (kind) {
    case 0:
        return n->i16 ? n->cbf_dc : 0;
    case 1: {
        const int b8 = ((r >> 3) & 1) * 2 + ((r >> 1) & 1);
        if (!((n->cbp >> b8) & 1)) return 0;
        if (n->t8) return 1;  // an 8x8 block's coded_block_flag is inferred to be 1
        return (n->cbf_luma >> r) & 1;
    }
    case 2:
        return (n->cbp >> 4) ? (n->cbf_cdc >> cc) & 1 : 0;
    default:
        return (n->cbp >> 4) == 2 ? (n->cbf_cac[cc] >> r) & 1 : 0;
    }
}

// residual_block_cabac (7.3.5.3.3) of block category cat into coef (maxNumCoeff entries, scan
// order); with_cbf: coded_block_flag present, cbf_inc its ctxIdxInc.  Returns 1 when coded.
int cabac_block(Cabac &c, int cat, int max_num, bool with_cbf, int cbf_inc, int16_t *coef) {
    static const int kCbfOff[5] = {0, 4, 8, 12, 16};
    static const int kSigOff[5] = {0, 15, 29, 44, 47};
    static const int kAbsOff[5] = {0, 10, 20, 30, 39};
    for (int i = 0; i < max_num; i++) coef[i] = 0;
    if (with_cbf && !c.decision(85 + kCbfOff[cat] + cbf_inc)) return 0;
    const int sig_base = cat == 5 ? 402 : 105 + kSigOff[cat], last_base = cat == 5 ? 417 : 166 + kSigOff[cat];
    const int abs_base = cat == 5 ? 426 : 227 + kAbsOff[cat];
    int idx[64], n = 0, i = 0;
    for (; i < max_num - 1; i++) {
        const int si = cat == 5 ? kSig8x8[i] : cat == 3 ? std::min(i, 2) : i;
        if (c.decision(sig_base + si)) {
            idx[n++] = i;
            const int li = cat == 5 ? kLast8x8[i] : cat == 3 ? std::min(i, 2) : i;
            if (c.decision(last_base + li)) break;
        }
    }
    if (i == max_num - 1) idx[n++] = i;
    int eq1 = 0, gt1 = 0;
    for (int k = n - 1; k >= 0; k--) {
        int v;
        if (!c.decision(abs_base + (gt1 ? 0 : std::min(4, 1 + eq1)))) {
            v = 1;
            eq1++;
        } else {
            const int ctx = abs_base + 5 + std::min(4 - (cat == 3 ? 1 : 0), gt1);
            v = 2;
            while (v < 15 && c.decision(ctx)) v++;
            if (v >= 15) {  // UEG0 suffix
                int j = 0;
                while (c.bypass()) {
                    if (++j > 24) return -1;
                }
                int s = 1;
                while (j--) s = 2 * s + c.bypass();
                v = s + 14;
            }
            gt1++;
        }
        coef[idx[k]] = (int16_t)(c.bypass() ? -v : v);
    }
    return 1;
}

// macroblock_layer (7.3.5) with CABAC
int cabac_mb(CabacPic &cp, PicState &st, int mb, int &qp, bool &prev_qp_nz, h264g_mb &m, int16_t (&mvd)[2][16][2],
             uint8_t (&isyn)[16][2]) {
    Cabac &c = cp.c;
    const Slice &sl = *st.sl;
    const Pps &pps = *st.pps;
    CMb &cur = cp.mbs[mb];
    memset(&cur, 0, sizeof(cur));
    memset(cur.ref, -1, sizeof(cur.ref));
    memset(&m, 0, sizeof(m));
    memset(mvd, 0, sizeof(mvd));
    for (int l = 0; l < 2; l++)
        for (int q = 0; q < 4; q++) m.ref_idx[l][q] = -1;
    m.level_offset = (uint32_t)st.levels->size();
    bool intra = sl.slice_type == H264G_SLICE_I;
    int mb_type;
    if (intra) mb_type = cabac_mb_type_intra(cp, mb, 3, true);
    else if (sl.slice_type == H264G_SLICE_P) mb_type = cabac_mb_type_p(cp, mb, intra);
    else mb_type = cabac_mb_type_b(cp, mb, intra);
    cur.avail = 1;
    cur.intra = intra;
    if (intra && mb_type == 25) {  // I_PCM: the samples follow byte-aligned, then the engine restarts
        m.mb_type = H264G_MB_IPCM;
        m.qp_y = (uint8_t)qp;
        cur.pcm = 1;
        cur.cbp = 0x2f;
        Bits &bs = *c.bs;
        while (!bs.aligned()) bs.bit();
        for (int i = 0; i < 384; i++) st.levels->push_back((int16_t)bs.u(8));
        c.init_engine();
        prev_qp_nz = false;
        return bs.over ? h264g_set_error(H264G_E_ARG, "truncated I_PCM") : 0;
    }
    bool t8 = false, i16 = false, no_sub_lt8 = true;
    int cbp = 0;
    const bool is_b = sl.slice_type == H264G_SLICE_B;
    const int nref[2] = {sl.num_ref_idx[0], sl.num_ref_idx[1]};
    auto set_ref = [&](int l, int q, int v) {  // quadrant q of list l
        m.ref_idx[l][q] = (int8_t)v;
        for (int k = 0; k < 4; k++) cur.ref[l][((q >> 1) * 2 + (k >> 1)) * 4 + (q & 1) * 2 + (k & 1)] = (int8_t)v;
    };
    auto put_amvd = [&](int l, int bx, int by, int w, int h, int dx, int dy) {
        for (int y = by; y < by + h; y++)
            for (int x = bx; x < bx + w; x++) {
                cur.amvd[l][y * 4 + x][0] = (uint8_t)std::min(255, std::abs(dx));
                cur.amvd[l][y * 4 + x][1] = (uint8_t)std::min(255, std::abs(dy));
            }
    };
    auto read_mvd = [&](int l, int bx, int by, int w, int h) -> int {
        const int dx = cabac_mvd(cp, mb, cur, l, bx, by, 0);
        const int dy = cabac_mvd(cp, mb, cur, l, bx, by, 1);
        if (dx < -32768 || dx > 32767 || dy < -32768 || dy > 32767) return h264g_set_error(H264G_E_ARG, "mvd out of range");
        mvd[l][by * 4 + bx][0] = (int16_t)dx;
        mvd[l][by * 4 + bx][1] = (int16_t)dy;
        put_amvd(l, bx, by, w, h, dx, dy);
        return 0;
    };
    if (intra) {
        if (mb_type == 0) {
            if (pps.transform_8x8_mode) {
                const CMb *a = cp.mb_at(mb, -1, 0), *b = cp.mb_at(mb, 0, -1);
                t8 = c.decision(399 + (a && a->t8) + (b && b->t8));
            }
            const int n = t8 ? 4 : 16;
            for (int k = 0; k < n; k++) {
                isyn[k][0] = (uint8_t)c.decision(68);
                isyn[k][1] = 0;
                if (!isyn[k][0]) {  // rem_intra_pred_mode: 3 bins, least significant first
                    int v = c.decision(69);
                    v |= c.decision(69) << 1;
                    v |= c.decision(69) << 2;
                    isyn[k][1] = (uint8_t)v;
                }
            }
            m.mb_type = t8 ? H264G_MB_I8x8 : H264G_MB_I4x4;
            cur.nxn = 1;
        } else {
            i16 = true;
            m.mb_type = H264G_MB_I16x16;
            m.intra16x16_pred_mode = (uint8_t)((mb_type - 1) % 4);
            cbp = (((mb_type - 1) / 4) % 3) << 4 | (mb_type >= 13 ? 15 : 0);
            cur.i16 = 1;
        }
        {  // intra_chroma_pred_mode (9.3.3.1.1.8)
            const CMb *a = cp.mb_at(mb, -1, 0), *b = cp.mb_at(mb, 0, -1);
            auto term = [](const CMb *n) { return n && n->intra && !n->pcm && n->chroma_mode != 0; };
            int v = 0;
            if (c.decision(64 + term(a) + term(b))) {
                v = 1;
                while (v < 3 && c.decision(64 + 3)) v++;
            }
            m.intra_chroma_pred_mode = (uint8_t)v;
            cur.chroma_mode = (uint8_t)v;
        }
    } else if (!is_b || mb_type != 0) {
        const InterType &it = is_b ? kBType[mb_type] : kPType[mb_type];
        m.mb_type = is_b ? H264G_MB_B : H264G_MB_P;
        m.part_shape = (uint8_t)it.shape;
        if (it.shape == H264G_PART_8x8) {
            int sub[4];
            for (int q = 0; q < 4; q++) {
                sub[q] = is_b ? cabac_sub_b(c) : cabac_sub_p(c);
                m.sub_part_shape[q] = (uint8_t)(is_b ? kBSubShape[sub[q]] : kPSubShape[sub[q]]);
                if (is_b && sub[q] == 0) {
                    m.flags |= H264G_MBF_DIRECT(q);
                    for (int k = 0; k < 4; k++) cur.direct[((q >> 1) * 2 + (k >> 1)) * 4 + (q & 1) * 2 + (k & 1)] = 1;
                    if (!st.sps->direct_8x8_inference) no_sub_lt8 = false;
                } else if (sub_parts(m.sub_part_shape[q]) > 1) {
                    no_sub_lt8 = false;
                }
            }
            auto pred = [&](int q) { return is_b ? kBSubPred[sub[q]] : 1; };
            for (int l = 0; l < 2; l++)
                for (int q = 0; q < 4; q++) {
                    if (!(pred(q) & (1 << l))) continue;
                    int v = 0;
                    if (nref[l] > 1) v = cabac_ref_idx(cp, mb, cur, l, (q & 1) * 2, (q >> 1) * 2, is_b);
                    if (v < 0 || v >= nref[l]) return h264g_set_error(H264G_E_ARG, "ref_idx out of range");
                    set_ref(l, q, v);
                }
            for (int l = 0; l < 2; l++)
                for (int q = 0; q < 4; q++) {
                    if (!(pred(q) & (1 << l))) continue;
                    const int shape = m.sub_part_shape[q];
                    const int w = shape == H264G_SUB_8x8 || shape == H264G_SUB_8x4 ? 2 : 1;
                    const int h = shape == H264G_SUB_8x8 || shape == H264G_SUB_4x8 ? 2 : 1;
                    for (int k = 0; k < sub_parts(shape); k++) {
                        int x, y;
                        sub_origin(shape, k, x, y);
                        if (int rc = read_mvd(l, (q & 1) * 2 + x, (q >> 1) * 2 + y, w, h)) return rc;
                    }
                }
        } else {
            const int np = it.shape == H264G_PART_16x16 ? 1 : 2;
            auto origin = [&](int k, int &bx, int &by, int &w, int &h) {
                bx = it.shape == H264G_PART_8x16 ? 2 * k : 0;
                by = it.shape == H264G_PART_16x8 ? 2 * k : 0;
                w = it.shape == H264G_PART_8x16 ? 2 : 4;
                h = it.shape == H264G_PART_16x8 ? 2 : 4;
            };
            for (int l = 0; l < 2; l++)
                for (int k = 0; k < np; k++) {
                    if (!(it.pred[k] & (1 << l))) continue;
                    int bx, by, w, h;
                    origin(k, bx, by, w, h);
                    int v = 0;
                    if (nref[l] > 1) v = cabac_ref_idx(cp, mb, cur, l, bx, by, is_b);
                    if (v < 0 || v >= nref[l]) return h264g_set_error(H264G_E_ARG, "ref_idx out of range");
                    for (int q = 0; q < 4; q++) {
                        const int qx = (q & 1) * 2, qy = (q >> 1) * 2;
                        if (qx >= bx && qx < bx + w && qy >= by && qy < by + h) set_ref(l, q, v);
                    }
                }
            for (int l = 0; l < 2; l++)
                for (int k = 0; k < np; k++) {
                    if (!(it.pred[k] & (1 << l))) continue;
                    int bx, by, w, h;
                    origin(k, bx, by, w, h);
                    if (int rc = read_mvd(l, bx, by, w, h)) return rc;
                }
        }
    } else {  // B_Direct_16x16
        m.mb_type = H264G_MB_B;
        m.part_shape = H264G_PART_8x8;
        m.flags |= H264G_MBF_DIRECT_ALL;
        cur.bdirect16 = 1;
        memset(cur.direct, 1, sizeof(cur.direct));
        if (!st.sps->direct_8x8_inference) no_sub_lt8 = false;
    }
    if (!i16) {  // coded_block_pattern (9.3.2.6, 9.3.3.1.1.4)
        const CMb *a = cp.mb_at(mb, -1, 0), *b = cp.mb_at(mb, 0, -1);
        for (int b8 = 0; b8 < 4; b8++) {
            const int bx = b8 & 1, by = b8 >> 1;
            int ta, tb;
            if (bx) ta = !((cbp >> (b8 - 1)) & 1);
            else ta = a && !a->pcm && (a->skip || !((a->cbp >> (b8 + 1)) & 1));
            if (by) tb = !((cbp >> (b8 - 2)) & 1);
            else tb = b && !b->pcm && (b->skip || !((b->cbp >> (b8 + 2)) & 1));
            cbp |= c.decision(73 + ta + 2 * tb) << b8;
        }
        auto cterm = [](const CMb *n, int need) { return n && (n->pcm || (!n->skip && (n->cbp >> 4) >= need)); };
        if (c.decision(77 + cterm(a, 1) + 2 * cterm(b, 1))) cbp |= (1 + c.decision(77 + 4 + cterm(a, 2) + 2 * cterm(b, 2))) << 4;
        const bool nxn = m.mb_type == H264G_MB_I4x4 || m.mb_type == H264G_MB_I8x8;
        if ((cbp & 15) && pps.transform_8x8_mode && !nxn && no_sub_lt8 && !intra) {
            const CMb *aa = cp.mb_at(mb, -1, 0), *bb = cp.mb_at(mb, 0, -1);
            t8 = c.decision(399 + (aa && aa->t8) + (bb && bb->t8));
        }
    }
    if (t8) m.flags |= H264G_MBF_T8x8;
    cur.t8 = t8;
    cur.cbp = (uint8_t)cbp;
    m.cbp = (uint8_t)cbp;
    if ((cbp & 15) || (cbp >> 4) || i16) {
        // mb_qp_delta (9.3.3.1.1.5): the previous MB in decoding order had a non-zero one
        int k = 0;
        if (c.decision(60 + (prev_qp_nz ? 1 : 0))) {
            k = 1;
            int ctx = 60 + 2;
            while (c.decision(ctx)) {
                if (++k > 104) return h264g_set_error(H264G_E_ARG, "mb_qp_delta out of range");
                ctx = 60 + 3;
            }
        }
        const int dq = std::min(25, std::max(-26, (k & 1) ? (k + 1) / 2 : -(k / 2)));
        qp = (qp + dq + 52) % 52;
        prev_qp_nz = dq != 0;
        cur.qp_delta_nz = dq != 0;
        // residual (7.3.5.3) with residual_block_cabac
        Residual r;
        memset(&r, 0, sizeof(r));
        const CMb *ma = cp.mb_at(mb, -1, 0), *mbb = cp.mb_at(mb, 0, -1);
        if (i16) {
            const int inc = cbf_term(ma, true, 0, 0, 0) + 2 * cbf_term(mbb, true, 0, 0, 0);
            cur.cbf_dc = (uint8_t)cabac_block(c, 0, 16, true, inc, r.dc);
        }
        for (int q = 0; q < 4; q++) {
            if (!((cbp >> q) & 1)) continue;
            if (t8) {
                const int got = cabac_block(c, 5, 64, false, 0, r.luma8[q]);
                if (got < 0) return h264g_set_error(H264G_E_ARG, "invalid coeff_abs_level_minus1");
                continue;
            }
            for (int k2 = 0; k2 < 4; k2++) {
                const int blk = 4 * q + k2, rr = blk_raster(blk), bx = rr & 3, by = rr >> 2;
                int ra, rb;
                const CMb *na = blk_at(cp, mb, cur, bx - 1, by, ra), *nb = blk_at(cp, mb, cur, bx, by - 1, rb);
                // inside the current MB the neighbour block is one of its own: its flag so far
                const int ta = na == &cur ? (cur.cbf_luma >> ra) & 1 : cbf_term(na, intra, 1, ra, 0);
                const int tb = nb == &cur ? (cur.cbf_luma >> rb) & 1 : cbf_term(nb, intra, 1, rb, 0);
                int16_t co[16];
                const int got = i16 ? cabac_block(c, 1, 15, true, ta + 2 * tb, co) : cabac_block(c, 2, 16, true, ta + 2 * tb, co);
                if (got < 0) return h264g_set_error(H264G_E_ARG, "invalid coeff_abs_level_minus1");
                memcpy(r.luma[blk], co, sizeof(co));
                if (got) cur.cbf_luma |= (uint16_t)(1u << rr);
            }
        }
        if (cbp >> 4) {
            for (int cc = 0; cc < 2; cc++) {
                const int inc = cbf_term(ma, intra, 2, 0, cc) + 2 * cbf_term(mbb, intra, 2, 0, cc);
                const int got = cabac_block(c, 3, 4, true, inc, r.cdc[cc]);
                if (got < 0) return h264g_set_error(H264G_E_ARG, "invalid coeff_abs_level_minus1");
                if (got) cur.cbf_cdc |= (uint8_t)(1u << cc);
            }
        }
        if ((cbp >> 4) == 2) {
            for (int cc = 0; cc < 2; cc++)
                for (int b4 = 0; b4 < 4; b4++) {
                    const int bx = b4 & 1, by = b4 >> 1;
                    int ta, tb;
                    if (bx) ta = (cur.cbf_cac[cc] >> (b4 - 1)) & 1;
                    else ta = cbf_term(ma, intra, 3, b4 + 1, cc);
                    if (by) tb = (cur.cbf_cac[cc] >> (b4 - 2)) & 1;
                    else tb = cbf_term(mbb, intra, 3, b4 + 2, cc);
                    int16_t co[16];
                    const int got = cabac_block(c, 4, 15, true, ta + 2 * tb, co);
                    if (got < 0) return h264g_set_error(H264G_E_ARG, "invalid coeff_abs_level_minus1");
                    memcpy(r.cac[cc][b4], co, 15 * sizeof(int16_t));
                    if (got) cur.cbf_cac[cc] |= (uint8_t)(1u << b4);
                }
        }
        emit_levels(r, i16, t8, cbp, *st.levels);
    } else {
        prev_qp_nz = false;
    }
    m.qp_y = (uint8_t)qp;
    return c.bs->over ? h264g_set_error(H264G_E_ARG, "truncated macroblock") : 0;
}

// slice_data() (7.3.4) with CABAC: cabac_alignment_one_bit, then mb_skip_flag /
// macroblock_layer / end_of_slice_flag per MB
int parse_slice_data_cabac(Bits &bs, PicState &st, int &mb) {
    const Slice &sl = *st.sl;
    const int n_mbs = st.W * st.H;
    while (!bs.aligned()) {
        if (!bs.bit()) return h264g_set_error(H264G_E_ARG, "cabac_alignment_one_bit is 0");
    }
    CabacPic cp;
    cp.W = st.W;
    cp.H = st.H;
    cp.mbs.assign((size_t)n_mbs, CMb());
    cp.c.bs = &bs;
    cp.c.init_contexts(sl.slice_type, sl.cabac_init_idc, sl.slice_qp);
    cp.c.init_engine();
    int qp = sl.slice_qp;
    bool prev_qp_nz = false;
    int16_t mvd[2][16][2];
    uint8_t isyn[16][2];
    while (mb < n_mbs) {
        st.mark = st.levels->size();
        h264g_mb &m = st.mbs[mb];
        if (sl.slice_type != H264G_SLICE_I && cabac_mb_skip(cp, mb, sl.slice_type)) {
            skip_record(m, sl.slice_type, qp, (uint32_t)st.levels->size());
            CMb &cm = cp.mbs[mb];
            memset(&cm, 0, sizeof(cm));
            memset(cm.ref, -1, sizeof(cm.ref));
            cm.avail = 1;
            cm.skip = 1;
            if (sl.slice_type == H264G_SLICE_B) {
                cm.bdirect16 = 1;
                memset(cm.direct, 1, sizeof(cm.direct));
            } else {
                memset(cm.ref[0], 0, sizeof(cm.ref[0]));
            }
            prev_qp_nz = false;
            if (int rc = finish_mb(st, mb, m, nullptr, nullptr)) return mb_error(st, mb, rc);
        } else {
            if (int rc = cabac_mb(cp, st, mb, qp, prev_qp_nz, m, mvd, isyn)) return mb_error(st, mb, rc);
            if (int rc = finish_mb(st, mb, m, &mvd[0][0][0], &isyn[0][0])) return mb_error(st, mb, rc);
        }
        mb++;
        if (cp.c.terminate()) break;  // end_of_slice_flag
        if (bs.over) return h264g_set_error(H264G_E_ARG, "picture %d: slice data truncated", st.picture);
    }
    return 0;
}

}  // namespace

// ------------------------------------------------------------------ DPB and pictures
namespace {

using Frame = h264g_parser::Frame;

int max_frame_num(const Sps &s) { return 1 << s.log2_max_frame_num; }

// FrameNumWrap (8-27) of a short-term frame relative to the current frame_num
int frame_num_wrap(const Frame &f, int cur_frame_num, int max_fn) {
    return f.frame_num > cur_frame_num ? f.frame_num - max_fn : f.frame_num;
}

// 8.2.1: PicOrderCnt of the current frame (types 0 and 2)
int decode_poc(h264g_parser *p, const Sps &sps, const Slice &sl) {
    if (sps.poc_type == 0) {
        if (sl.idr) {
            p->prev_poc_msb = 0;
            p->prev_poc_lsb = 0;
        } else if (p->prev_mmco5) {
            p->prev_poc_msb = 0;
            p->prev_poc_lsb = p->prev_ref_top_poc;
        }
        const int max_lsb = 1 << sps.log2_max_poc_lsb;
        int msb;
        if (sl.poc_lsb < p->prev_poc_lsb && p->prev_poc_lsb - sl.poc_lsb >= max_lsb / 2) msb = p->prev_poc_msb + max_lsb;
        else if (sl.poc_lsb > p->prev_poc_lsb && sl.poc_lsb - p->prev_poc_lsb > max_lsb / 2) msb = p->prev_poc_msb - max_lsb;
        else msb = p->prev_poc_msb;
        const int top = msb + sl.poc_lsb, bottom = top + sl.delta_poc_bottom;
        if (sl.nal_ref_idc) {  // prevPicOrderCntMsb / Lsb come from the previous reference picture
            p->prev_poc_msb = msb;
            p->prev_poc_lsb = sl.poc_lsb;
        }
        return std::min(top, bottom);
    }
    if (sps.poc_type == 1) {  // 8.2.1.2, as the reference derives it (H264RefPicList.cpp:118-215)
        // prevFrameNumOffset and the frame_num of the previous picture in decoding order
        const int prev_off = p->last_pic_mmco5 ? 0 : p->prev_frame_num_offset;
        const int offset = sl.idr ? 0 : (p->prev_raw_frame_num > sl.frame_num ? prev_off + max_frame_num(sps) : prev_off);
        int abs_fn = sps.n_cycle ? offset + sl.frame_num : 0;
        if (!sl.nal_ref_idc && abs_fn > 0) abs_fn--;
        int expected = 0;
        if (abs_fn > 0) {
            const int cycle = (abs_fn - 1) / sps.n_cycle, in_cycle = (abs_fn - 1) % sps.n_cycle;
            expected = cycle * sps.expected_delta_cycle;
            for (int i = 0; i <= in_cycle; i++) expected += sps.offset_ref[i];
        }
        if (!sl.nal_ref_idc) expected += sps.offset_non_ref;
        const int top = expected + sl.delta_poc[0], bottom = top + sps.offset_top_to_bottom + sl.delta_poc[1];
        p->prev_frame_num_offset = offset;
        return std::min(top, bottom);
    }
    // type 2
    int offset;
    if (sl.idr) offset = 0;
    else if (p->prev_frame_num > sl.frame_num) offset = p->prev_frame_num_offset + max_frame_num(sps);
    else offset = p->prev_frame_num_offset;
    if (p->prev_mmco5 && !sl.idr) offset = p->prev_frame_num > sl.frame_num ? max_frame_num(sps) : 0;
    const int poc = sl.idr ? 0 : (sl.nal_ref_idc ? 2 * (offset + sl.frame_num) : 2 * (offset + sl.frame_num) - 1);
    p->prev_frame_num_offset = offset;
    return poc;
}

// 8.2.4.2.1 / 8.2.4.2.3 / 8.2.4.2.4 (frames): initial lists of DPB indices
void init_lists(h264g_parser *p, const Sps &sps, const Slice &sl, int cur_poc, std::vector<int> lists[2]) {
    const int max_fn = max_frame_num(sps);
    std::vector<int> st, lt;
    for (int i = 0; i < (int)p->dpb.size(); i++) {
        if (p->dpb[i].short_ref) st.push_back(i);
        else if (p->dpb[i].long_ref) lt.push_back(i);
    }
    std::sort(lt.begin(), lt.end(), [&](int a, int b) { return p->dpb[a].long_term_idx < p->dpb[b].long_term_idx; });
    lists[0].clear();
    lists[1].clear();
    if (sl.slice_type == H264G_SLICE_P) {
        std::sort(st.begin(), st.end(), [&](int a, int b) {
            return frame_num_wrap(p->dpb[a], sl.frame_num, max_fn) > frame_num_wrap(p->dpb[b], sl.frame_num, max_fn);
        });
        lists[0] = st;
        lists[0].insert(lists[0].end(), lt.begin(), lt.end());
    } else if (sl.slice_type == H264G_SLICE_B) {
        std::vector<int> before, after;
        for (int i : st) (p->dpb[i].poc < cur_poc ? before : after).push_back(i);
        std::sort(before.begin(), before.end(), [&](int a, int b) { return p->dpb[a].poc > p->dpb[b].poc; });
        std::sort(after.begin(), after.end(), [&](int a, int b) { return p->dpb[a].poc < p->dpb[b].poc; });
        lists[0] = before;
        lists[0].insert(lists[0].end(), after.begin(), after.end());
        lists[0].insert(lists[0].end(), lt.begin(), lt.end());
        lists[1] = after;
        lists[1].insert(lists[1].end(), before.begin(), before.end());
        lists[1].insert(lists[1].end(), lt.begin(), lt.end());
        if (lists[1].size() > 1 && lists[0] == lists[1]) std::swap(lists[1][0], lists[1][1]);
    }
}

// 8.2.4.3: modification of a list (entries are DPB indices, -1 = no reference picture)
int modify_list(h264g_parser *p, const Sps &sps, const Slice &sl, int l, std::vector<int> &list) {
    const int max_fn = max_frame_num(sps), n = sl.num_ref_idx[l];
    list.resize(n, -1);
    int pred = sl.frame_num;
    int ref_idx = 0;
    for (const auto &md : sl.mods[l]) {
        int target = -1;
        if (md.first < 2) {
            const int abs_diff = md.second + 1;
            int no_wrap = md.first == 0 ? pred - abs_diff : pred + abs_diff;
            if (no_wrap < 0) no_wrap += max_fn;
            if (no_wrap >= max_fn) no_wrap -= max_fn;
            pred = no_wrap;
            const int pic_num = no_wrap > sl.frame_num ? no_wrap - max_fn : no_wrap;
            for (int i = 0; i < (int)p->dpb.size(); i++)
                if (p->dpb[i].short_ref && frame_num_wrap(p->dpb[i], sl.frame_num, max_fn) == pic_num) target = i;
        } else {
            for (int i = 0; i < (int)p->dpb.size(); i++)
                if (p->dpb[i].long_ref && p->dpb[i].long_term_idx == md.second) target = i;
        }
        if (target < 0) return h264g_set_error(H264G_E_ARG, "reference list modification names a missing picture");
        if (ref_idx >= n) return h264g_set_error(H264G_E_ARG, "too many reference list modifications");
        list.insert(list.begin() + ref_idx, target);
        ref_idx++;
        int w = ref_idx;
        for (int r = ref_idx; r < (int)list.size(); r++)
            if (list[r] != target) list[w++] = list[r];
        list.resize(w);
        list.resize(n, -1);
    }
    list.resize(n, -1);
    return 0;
}

// 8.2.5: marking after decoding the current picture (which becomes DPB entry `cur`)
void mark(h264g_parser *p, const Sps &sps, const Slice &sl, Frame &cur) {
    const int max_fn = max_frame_num(sps);
    p->prev_mmco5 = false;
    if (sl.idr) {
        for (Frame &f : p->dpb) f.short_ref = f.long_ref = false;
        if (sl.long_term_reference) {
            cur.long_ref = true;
            cur.long_term_idx = 0;
            p->max_long_term_idx = 0;
        } else {
            cur.short_ref = true;
            p->max_long_term_idx = -1;
        }
        return;
    }
    bool cur_long = false;
    if (sl.adaptive_marking) {
        for (const Mmco &m : sl.mmco) {
            switch (m.op) {
            case 1: {
                const int pn = sl.frame_num - m.diff;
                for (Frame &f : p->dpb)
                    if (f.short_ref && frame_num_wrap(f, sl.frame_num, max_fn) == pn) f.short_ref = false;
                break;
            }
            case 2:
                for (Frame &f : p->dpb)
                    if (f.long_ref && f.long_term_idx == m.lt_pic_num) f.long_ref = false;
                break;
            case 3: {
                const int pn = sl.frame_num - m.diff;
                for (Frame &f : p->dpb)
                    if (f.long_ref && f.long_term_idx == m.lt_idx) f.long_ref = false;
                for (Frame &f : p->dpb)
                    if (f.short_ref && frame_num_wrap(f, sl.frame_num, max_fn) == pn) {
                        f.short_ref = false;
                        f.long_ref = true;
                        f.long_term_idx = m.lt_idx;
                    }
                break;
            }
            case 4:
                p->max_long_term_idx = m.max_lt_idx_plus1 - 1;
                for (Frame &f : p->dpb)
                    if (f.long_ref && f.long_term_idx > p->max_long_term_idx) f.long_ref = false;
                break;
            case 5:
                for (Frame &f : p->dpb) f.short_ref = f.long_ref = false;
                p->max_long_term_idx = -1;
                p->prev_mmco5 = true;
                break;
            case 6:
                for (Frame &f : p->dpb)
                    if (f.long_ref && f.long_term_idx == m.lt_idx) f.long_ref = false;
                cur.long_ref = true;
                cur.long_term_idx = m.lt_idx;
                cur_long = true;
                break;
            }
        }
    } else {  // sliding window (8.2.5.3)
        int n_short = 0, n_long = 0;
        for (const Frame &f : p->dpb) {
            n_short += f.short_ref;
            n_long += f.long_ref;
        }
        if (n_short + n_long >= std::max(sps.max_num_ref_frames, 1) && n_short > 0) {
            int oldest = -1;
            for (int i = 0; i < (int)p->dpb.size(); i++)
                if (p->dpb[i].short_ref &&
                    (oldest < 0 || frame_num_wrap(p->dpb[i], sl.frame_num, max_fn) <
                                       frame_num_wrap(p->dpb[oldest], sl.frame_num, max_fn)))
                    oldest = i;
            p->dpb[oldest].short_ref = false;
        }
    }
    if (!cur_long) cur.short_ref = true;
}

// the next NAL unit: [begin, end) of its bytes (header included), or false at the end
bool next_nal(h264g_parser *p, size_t &begin, size_t &end) {
    const std::vector<uint8_t> &d = p->data;
    size_t i = p->pos;
    // start code prefix 0x000001
    while (i + 3 <= d.size() && !(d[i] == 0 && d[i + 1] == 0 && d[i + 2] == 1)) i++;
    if (i + 3 > d.size()) {
        p->pos = d.size();
        return false;
    }
    begin = i + 3;
    size_t j = begin;
    while (j + 3 <= d.size() && !(d[j] == 0 && d[j + 1] == 0 && (d[j + 2] == 1 || (d[j + 2] == 0 && j + 3 < d.size() && d[j + 3] == 1))))
        j++;
    end = j + 3 <= d.size() ? j : d.size();
    while (end > begin && d[end - 1] == 0) end--;  // trailing_zero_8bits
    p->pos = j + 3 <= d.size() ? j : d.size();
    return true;
}

}  // namespace

extern "C" int32_t h264g_parser_create(const uint8_t *data, uint64_t len, h264g_parser **out) {
    if (!data || !out) return h264g_set_error(H264G_E_ARG, "NULL argument");
    *out = nullptr;
    h264g_parser *p = new (std::nothrow) h264g_parser();
    if (!p) return h264g_set_error(H264G_E_NOMEM, "parser");
    p->data.assign(data, data + len);
    // pictures whose slice data parses ahead, each on a thread of its own: the environment's
    // H264G_PARSE_THREADS (1: every picture parses in order on the calling thread), else
    // the host's hardware threads, at most 8
    const char *env = getenv("H264G_PARSE_THREADS");
    const int n = env && *env ? atoi(env) : (int)std::thread::hardware_concurrency();
    p->threads = std::max(1, std::min(n, env && *env ? 64 : 8));
    *out = p;
    return H264G_OK;
}

extern "C" int32_t h264g_parser_destroy(h264g_parser *p) {
    delete p;
    return H264G_OK;
}

namespace {

// The slice's reference lists as DPB indices: initialisation, then modification.
int slice_lists(h264g_parser *p, const Sps &sps, const Slice &sl, int poc, std::vector<int> lists[2]) {
    init_lists(p, sps, sl, poc, lists);
    for (int l = 0; l < (sl.slice_type == H264G_SLICE_B ? 2 : sl.slice_type == H264G_SLICE_P ? 1 : 0); l++) {
        if (int rc = modify_list(p, sps, sl, l, lists[l])) return rc;
    }
    return 0;
}

bool explicit_wp(const Pps &pps, const Slice &sl) {
    return (pps.weighted_pred && sl.slice_type == H264G_SLICE_P) || (pps.weighted_bipred_idc == 1 && sl.slice_type == H264G_SLICE_B);
}

// Fills the picture's descriptor from its first slice (except dst_slot).
// Default_4x4_Intra / _Inter and Default_8x8_Intra / _Inter (Table 7-3, 7-4), scan order.
const uint8_t kDefault4[2][16] = {{6, 13, 13, 20, 20, 20, 28, 28, 28, 28, 32, 32, 32, 37, 37, 42},
                                  {10, 14, 14, 20, 20, 20, 24, 24, 24, 24, 27, 27, 27, 30, 30, 34}};
const uint8_t kDefault8[2][64] = {
    {6,  10, 10, 13, 11, 13, 16, 16, 16, 16, 18, 18, 18, 18, 18, 23, 23, 23, 23, 23, 23, 25,
     25, 25, 25, 25, 25, 25, 27, 27, 27, 27, 27, 27, 27, 27, 29, 29, 29, 29, 29, 29, 29, 31,
     31, 31, 31, 31, 31, 33, 33, 33, 33, 33, 36, 36, 36, 36, 38, 38, 38, 40, 40, 42},
    {9,  13, 13, 15, 13, 15, 17, 17, 17, 17, 19, 19, 19, 19, 19, 21, 21, 21, 21, 21, 21, 22,
     22, 22, 22, 22, 22, 22, 24, 24, 24, 24, 24, 24, 24, 24, 25, 25, 25, 25, 25, 25, 25, 27,
     27, 27, 27, 27, 27, 28, 28, 28, 28, 28, 30, 30, 30, 30, 32, 32, 32, 33, 33, 35}};

// Raster position of each zig-zag scan position of an n x n block (Tables 8-12, 8-13):
// anti-diagonals in turn, odd ones walked down-left, even ones up-right.
template <int N>
struct ZigZag {
    uint8_t at[N * N];
    constexpr ZigZag() : at() {
        int k = 0;
        for (int d = 0; d < 2 * N - 1; d++)
            for (int t = 0; t <= d; t++) {
                const int i = d & 1 ? t : d - t, j = d - i;  // row i, column j
                if (i < N && j < N) at[k++] = (uint8_t)(i * N + j);
            }
    }
};
constexpr ZigZag<4> kZz4;
constexpr ZigZag<8> kZz8;

// The picture's scaling lists as the reference's set_scaling_lists_values leaves them
// (H264SliceHeader.cpp:928-1158): Table 7-2 fall-back rules A (SPS) then B (PPS), in raster
// order.  Quirk: a list the PPS sends explicitly is taken from the SPS's parsed list of
// that index (zero when the SPS did not send it), not from the PPS (:1093-1106, 1136-1149).
void scaling_desc(const Sps &sps, const Pps &pps, h264g_pic_desc &d) {
    const ScalingLists &a = sps.sl, &b = pps.sl;
    if (!a.present_matrix && !b.present_matrix) return;  // flat
    uint8_t l4[6][16] = {}, l8[2][64] = {};
    if (a.present_matrix) {
        for (int i = 0; i < 6; i++) {
            const uint8_t *src = !a.present[i] ? (i == 0 || i == 3 ? kDefault4[i / 3] : l4[i - 1])
                               : a.use_default[i] ? kDefault4[i / 3] : a.l4[i];
            memcpy(l4[i], src, 16);
        }
        for (int i = 0; i < 2; i++)
            memcpy(l8[i], !a.present[6 + i] || a.use_default[6 + i] ? kDefault8[i] : a.l8[i], 64);
    }
    if (b.present_matrix) {
        for (int i = 0; i < 6; i++) {
            if (!b.present[i]) {
                if (i != 0 && i != 3) memcpy(l4[i], l4[i - 1], 16);
                else if (!a.present_matrix) memcpy(l4[i], kDefault4[i / 3], 16);
            } else {
                memcpy(l4[i], b.use_default[i] ? kDefault4[i / 3] : a.l4[i], 16);
            }
        }
        for (int i = 0; i < 2; i++) {
            if (!b.present[6 + i]) {
                if (!a.present_matrix) memcpy(l8[i], kDefault8[i], 64);
            } else {
                memcpy(l8[i], b.use_default[6 + i] ? kDefault8[i] : a.l8[i], 64);
            }
        }
    }
    d.scaling_matrix = 1;
    for (int i = 0; i < 6; i++)
        for (int k = 0; k < 16; k++) d.weight_scale4x4[i][kZz4.at[k]] = l4[i][k];
    for (int i = 0; i < 2; i++)
        for (int k = 0; k < 64; k++) d.weight_scale8x8[i][kZz8.at[k]] = l8[i][k];
}

void fill_desc(const h264g_parser *p, const Sps &sps, const Pps &pps, const Slice &sl, int poc,
               const std::vector<int> lists[2], h264g_pic_desc &d) {
    memset(&d, 0, sizeof(d));
    scaling_desc(sps, pps, d);
    d.slice_type = sl.slice_type;
    d.structure = H264G_FRAME;
    d.constrained_intra_pred_flag = pps.constrained_intra_pred;
    d.disable_deblocking_filter_idc = sl.disable_dbk;
    d.filter_offset_a = sl.alpha_div2 * 2;
    d.filter_offset_b = sl.beta_div2 * 2;
    d.chroma_qp_index_offset = pps.chroma_qp_index_offset;
    d.second_chroma_qp_index_offset = pps.second_chroma_qp_index_offset;
    d.weighted_pred_flag = pps.weighted_pred;
    d.weighted_bipred_idc = pps.weighted_bipred_idc;
    d.luma_log2_weight_denom = sl.luma_log2_wd;
    d.chroma_log2_weight_denom = sl.chroma_log2_wd;
    d.poc = poc;
    for (int l = 0; l < 2; l++) {
        d.num_ref_idx_active[l] = sl.num_ref_idx[l];
        d.luma_weight_flag_last[l] = sl.luma_flag_last[l];
        for (int r = 0; r < 32; r++) {
            d.ref_slot[l][r] = -1;
            if (r < (int)lists[l].size() && lists[l][r] >= 0) {
                const Frame &f = p->dpb[lists[l][r]];
                d.ref_slot[l][r] = f.slot;
                d.ref_poc[l][r] = f.poc;
                d.ref_long_term[l][r] = f.long_ref ? 1 : 0;
            }
            if (explicit_wp(pps, sl) && r < sl.num_ref_idx[l]) {
                d.luma_weight[l][r] = sl.lw[l][r];
                d.luma_offset[l][r] = sl.lo[l][r];
                for (int c = 0; c < 2; c++) {
                    d.chroma_weight[l][r][c] = sl.cw[l][r][c];
                    d.chroma_offset[l][r][c] = sl.co[l][r][c];
                }
            }
        }
    }
}

// Why a later slice of a picture cannot share the picture's descriptor, or null.
const char *slice_mismatch(const Pps &pps, const Slice &a, const Slice &b, const std::vector<int> la[2],
                           const std::vector<int> lb[2]) {
    if (a.slice_type != b.slice_type) return "slice types differ";
    if (a.frame_num != b.frame_num || a.pps_id != b.pps_id || a.idr != b.idr || (a.nal_ref_idc != 0) != (b.nal_ref_idc != 0))
        return "frame_num / PPS / IDR / reference flag differ";
    if (a.disable_dbk != b.disable_dbk || a.alpha_div2 != b.alpha_div2 || a.beta_div2 != b.beta_div2)
        return "deblocking parameters differ";
    for (int l = 0; l < 2; l++) {
        if (a.num_ref_idx[l] != b.num_ref_idx[l] || la[l] != lb[l]) return "reference lists differ";
        if (!explicit_wp(pps, a)) continue;
        if (a.luma_log2_wd != b.luma_log2_wd || a.chroma_log2_wd != b.chroma_log2_wd || a.luma_flag_last[l] != b.luma_flag_last[l])
            return "weight tables differ";
        for (int r = 0; r < a.num_ref_idx[l]; r++)
            if (a.lw[l][r] != b.lw[l][r] || a.lo[l][r] != b.lo[l][r] || memcmp(a.cw[l][r], b.cw[l][r], sizeof(a.cw[l][r])) ||
                memcmp(a.co[l][r], b.co[l][r], sizeof(a.co[l][r])))
                return "weight tables differ";
    }
    return nullptr;
}

}  // namespace

extern "C" const char *h264g_parser_damage(const h264g_parser *p) { return p ? p->damage : ""; }

namespace {

// Scans the next picture's NAL units: parameter sets update the tables, slice headers are
// parsed and each slice's RBSP is kept, up to the first NAL unit of the next picture.
// Null at the end of the stream; an error is kept in the picture's job, to be returned
// once the pictures before it have been handed out, and ends gathering.
std::unique_ptr<PicJob> gather(h264g_parser *p) {
    std::unique_ptr<PicJob> job(new PicJob());
    auto fail = [&](int32_t rc) {
        job->rc = rc;
        job->err = h264g_last_error();
        p->gather_end = true;
        return std::move(job);
    };
    std::vector<uint8_t> rbsp;
    size_t begin, end;
    const Pps *pps = nullptr;
    for (;;) {
        const size_t nal_pos = p->pos;
        if (!next_nal(p, begin, end)) break;
        if (end <= begin) continue;
        const uint8_t hdr = p->data[begin];
        const int nal_type = hdr & 31, nal_ref_idc = (hdr >> 5) & 3;
        const bool started = !job->slices.empty();
        if (nal_type != 1 && nal_type != 5) {
            if (nal_type >= 2 && nal_type <= 4) return fail(h264g_set_error(H264G_E_ARG, "data partitioning is not supported"));
            // parameter sets, SEI, access unit delimiters and sequence ends come before a
            // picture's first slice: one of them ends the picture being gathered (7.4.1.2.3)
            if (started && nal_type >= 6 && nal_type <= 11) {
                p->pos = nal_pos;
                break;
            }
            if (nal_type == 7 || nal_type == 8) {
                unescape(&p->data[begin + 1], end - begin - 1, rbsp);
                Bits bs;
                bs.d = rbsp.data();
                bs.nbits = rbsp.size() * 8;
                if (int rc = nal_type == 7 ? parse_sps(bs, p->sps) : parse_pps(bs, p->pps, p->sps)) return fail(rc);
            }
            continue;  // SEI, AUD, end of sequence / stream, filler ...
        }
        PicJob::SliceNal s;
        unescape(&p->data[begin + 1], end - begin - 1, s.rbsp);
        Bits bs;
        bs.d = s.rbsp.data();
        bs.nbits = s.rbsp.size() * 8;
        const Sps *ssps = nullptr;
        const Pps *spps = nullptr;
        if (int rc = parse_slice_header(bs, s.sl, nal_type, nal_ref_idc, p->sps, p->pps, &ssps, &spps)) return fail(rc);
        if (s.sl.redundant_pic_cnt) continue;
        if (started && s.sl.first_mb == 0) {  // the next picture's first slice
            p->pos = nal_pos;
            break;
        }
        if (s.sl.slice_type == H264G_SLICE_B && !s.sl.direct_spatial)
            return fail(h264g_set_error(H264G_E_ARG, "temporal direct prediction is not supported"));
        if (!started) {
            if (s.sl.first_mb != 0)
                return fail(h264g_set_error(H264G_E_ARG, "a picture starts with a slice at MB %d", s.sl.first_mb));
            job->sps = *ssps;
            job->pps = *spps;
            pps = spps;
        } else if (spps != pps) {
            return fail(h264g_set_error(H264G_E_ARG, "picture %d: slices refer to different PPSs", p->gathered));
        }
        s.data_bit = bs.pos;
        s.over = bs.over;
        job->slices.push_back(std::move(s));
    }
    if (job->slices.empty()) {
        p->gather_end = true;
        return nullptr;
    }
    p->gathered++;
    return job;
}

Bits slice_bits(const PicJob::SliceNal &s) {
    Bits bs;
    bs.d = s.rbsp.data();
    bs.nbits = s.rbsp.size() * 8;
    bs.pos = s.data_bit;
    bs.over = s.over;
    return bs;
}

// Parses a gathered picture's slice data ahead of its turn (on a thread of its own): the
// records and levels, with each MB's vector and mode syntax kept for the in-order pass.
// A damaged slice ends as the in-order parse ends it; anything else out of the ordinary
// (slices out of order, an error at the picture's end) is left to the in-order pass,
// which parses the picture again and reports it.
void parse_ahead_slices(PicJob *job, int picture) {
    const int W = job->sps.width_mbs, H = job->sps.height_map_units, n_mbs = W * H;
    job->mbs.assign((size_t)n_mbs, h264g_mb());
    job->levels.clear();
    // Without B slices nothing is predicted from another picture: the worker derives the
    // vectors and modes itself, into a motion field of its own in the state the in-order
    // pass starts a picture with.
    bool own = true;
    for (const PicJob::SliceNal &s : job->slices) own &= s.sl.slice_type != H264G_SLICE_B;
    if (own && (h264g_mvfield_create(W, H, &job->field) || h264g_mvfield_reset(job->field) ||
                h264g_mvfield_colocated(job->field, nullptr, 0)))
        own = false;
    h264g_mvfield *field = own ? job->field : nullptr;
    h264g_mb *mbs = job->mbs.data();
    if (!own) job->fin.resize((size_t)n_mbs);
    int covered = 0;
    uint32_t n_slices = 0;
    for (PicJob::SliceNal &s : job->slices) {
        if (s.sl.first_mb < covered || s.sl.first_mb >= n_mbs) return;
        for (; covered < s.sl.first_mb; covered++) {  // MBs the slices left out
            if (!own) undecoded_record(mbs, job->levels, covered, H264G_SLICE_NONE);
            else if (undecoded_mb(mbs, job->levels, field, covered, H264G_SLICE_NONE)) return;
        }
        Bits bs = slice_bits(s);
        PicState st{field, own ? nullptr : &job->fin, picture, &job->sps, &job->pps, &s.sl, W, H,
                    std::vector<MbNz>((size_t)n_mbs), mbs, &job->levels, n_slices};
        int mb = s.sl.first_mb;
        st.mark = job->levels.size();
        int rc = job->pps.cabac ? parse_slice_data_cabac(bs, st, mb) : parse_slice_data_cavlc(bs, st, mb);
        if (!rc && bs.over) rc = h264g_set_error(H264G_E_ARG, "picture %d: slice data truncated", picture);
        if (rc && mb >= n_mbs) return;
        if (rc) {  // the slice ends with the MB that failed (see assemble_slices)
            s.damaged_mb = mb;
            s.damage = h264g_last_error();
            if (!st.partial) {
                job->levels.resize(st.mark);
                if (!own) {
                    undecoded_record(mbs, job->levels, mb, n_slices);
                    job->fin[mb].has_mvd = 0;
                } else if (undecoded_mb(mbs, job->levels, field, mb, n_slices)) {
                    return;
                }
            }
            mb++;
        }
        s.mb_end = mb;
        covered = mb;
        n_slices++;
    }
    for (; covered < n_mbs; covered++) {
        if (!own) undecoded_record(mbs, job->levels, covered, H264G_SLICE_NONE);
        else if (undecoded_mb(mbs, job->levels, field, covered, H264G_SLICE_NONE)) return;
    }
    job->derived = own;
    job->ahead_ok = true;
}

void parse_ahead(PicJob *job, int picture) {
    try {
        parse_ahead_slices(job, picture);
    } catch (const std::exception &) {  // out of memory: the picture parses in order
        job->ahead_ok = false;
    }
}

constexpr int kParseAgain = 1;  // the syntax parsed ahead failed to derive: parse in order

// The picture's slices in order, into the records: slice data parsed now (damaged slices
// end where the reference abandons them), or, when it parsed ahead, the vectors and modes
// of its MBs derived in MB order.  MBs no slice covers are left undecoded.
int assemble_slices(h264g_parser *p, PicJob &job, bool ahead, int poc, const std::vector<int> lists[2],
                    h264g_parsed_picture *out) {
    const Sps &sps = job.sps;
    const Pps &pps = job.pps;
    const int W = sps.width_mbs, H = sps.height_map_units, n_mbs = W * H;
    int covered = 0, n_slices = 0, lost = 0, first_lost = -1;
    p->damage[0] = 0;
    if (!ahead) {
        job.mbs.assign((size_t)n_mbs, h264g_mb());
        job.levels.clear();
    }
    h264g_mb *mbs = job.mbs.data();
    for (size_t i = 0; i < job.slices.size(); i++) {
        PicJob::SliceNal &s = job.slices[i];
        const Slice &sl = s.sl;
        if (i > 0) {  // a later slice of the picture must fit the picture's descriptor
            std::vector<int> sl_lists[2];
            if (int rc = slice_lists(p, sps, sl, poc, sl_lists)) return rc;
            if (const char *why = slice_mismatch(pps, job.slices[0].sl, sl, lists, sl_lists))
                return h264g_set_error(H264G_E_ARG, "picture %d: its slices do not share one descriptor (%s)", p->pictures, why);
        }
        if (sl.first_mb < covered || sl.first_mb >= n_mbs)
            return h264g_set_error(H264G_E_ARG, "picture %d: slice at MB %d after MB %d", p->pictures, sl.first_mb, covered);
        for (; covered < sl.first_mb; covered++) {  // MBs the slices left out
            if (first_lost < 0) first_lost = covered;
            lost++;
            if (int rc = !ahead ? undecoded_mb(mbs, job.levels, p->field, covered, H264G_SLICE_NONE)
                         : job.derived ? 0 : h264g_mvfield_derive(p->field, (uint32_t)covered, &mbs[covered], nullptr))
                return rc;
        }
        int mb = sl.first_mb;
        if (ahead) {
            if (s.damaged_mb >= 0) {
                snprintf(p->damage, sizeof(p->damage), "%s", s.damage.c_str());
                if (first_lost < 0) first_lost = s.damaged_mb;
                lost++;
            }
            if (job.derived) mb = s.mb_end;
            for (; mb < s.mb_end; mb++) {
                const FinMb &f = job.fin[mb];
                if (derive_mb(p->field, pps.constrained_intra_pred, mb, mbs[mb], f.has_mvd ? &f.mvd[0][0][0] : nullptr,
                              &f.isyn[0][0]))
                    return kParseAgain;
            }
        } else {
            // slice data: neighbours of other slices are not available (fresh per-slice state)
            Bits bs = slice_bits(s);
            PicState st{p->field, nullptr, p->pictures, &sps, &pps, &sl, W, H, std::vector<MbNz>((size_t)n_mbs),
                        mbs, &job.levels, (uint32_t)n_slices};
            st.mark = job.levels.size();
            int rc = pps.cabac ? parse_slice_data_cabac(bs, st, mb) : parse_slice_data_cavlc(bs, st, mb);
            if (!rc && bs.over) rc = h264g_set_error(H264G_E_ARG, "picture %d: slice data truncated", p->pictures);
            if (rc && mb < n_mbs) {
                // Damaged slice data: the slice ends with the MB that failed, and decoding goes
                // on with the next slice, as the reference's does (H264VideoDecoder.cpp:135-136,
                // CONTINUE_IF_FAILED).  An MB whose residual failed (CAVLC) is reconstructed
                // from what parsed; any other failed MB stays undecoded.
                snprintf(p->damage, sizeof(p->damage), "%s", h264g_last_error());
                if (first_lost < 0) first_lost = mb;
                lost++;
                if (!st.partial) {
                    job.levels.resize(st.mark);
                    if (int rc2 = undecoded_mb(mbs, job.levels, p->field, mb, st.slice_num)) return rc2;
                }
                mb++;
            } else if (rc) {
                return rc;
            }
        }
        covered = mb;
        n_slices++;
    }
    for (; covered < n_mbs; covered++) {  // MBs after the last slice
        if (first_lost < 0) first_lost = covered;
        lost++;
        if (int rc = !ahead ? undecoded_mb(mbs, job.levels, p->field, covered, H264G_SLICE_NONE)
                     : job.derived ? 0 : h264g_mvfield_derive(p->field, (uint32_t)covered, &mbs[covered], nullptr))
            return rc;
    }
    out->n_slices = n_slices;
    out->lost_mbs = lost;
    out->first_lost_mb = first_lost;
    return 0;
}

// The motion field, fresh for the picture, with its co-located picture for direct prediction.
int start_field(h264g_parser *p, const Slice &sl, const std::vector<int> lists[2]) {
    h264g_mvfield_reset(p->field);
    if (sl.slice_type == H264G_SLICE_B) {
        if (lists[1].empty() || lists[1][0] < 0) return h264g_set_error(H264G_E_ARG, "B slice without RefPicList1[0]");
        const Frame &col = p->dpb[lists[1][0]];
        return h264g_mvfield_colocated(p->field, col.mbs.data(), col.short_ref ? 1 : 0);
    }
    return h264g_mvfield_colocated(p->field, nullptr, 0);
}

// A gathered picture in decoding order: POC, reference lists, descriptor and DPB slot,
// then its slices, then reference marking.
int assemble(h264g_parser *p, PicJob &job, h264g_parsed_picture *out) {
    const Sps &sps = job.sps;
    const Pps &pps = job.pps;
    const Slice &sl = job.slices[0].sl;
    const int W = sps.width_mbs, H = sps.height_map_units;
    if (p->field && (p->width_mbs != W || p->height_mbs != H))
        return h264g_set_error(H264G_E_ARG, "the picture size changes inside the stream");
    if (!p->field) {
        p->width_mbs = W;
        p->height_mbs = H;
        p->dpb_slots = std::max(sps.max_num_ref_frames, 1) + 1;
        if (h264g_mvfield_create(W, H, &p->field)) return h264g_set_error(H264G_E_NOMEM, "motion field");
    }
    if (sl.idr) {
        p->dpb.clear();
        p->prev_frame_num = 0;
    }
    // A gap in frame_num goes without the "non-existing" frames of 8.2.5.2: the
    // reference's Decoding_process_for_gaps_in_frame_num does nothing
    // (H264RefPicList.cpp:1598-1601), so the pictures that remain are the references.
    const int poc = decode_poc(p, sps, sl);
    std::vector<int> lists[2];
    if (int rc = slice_lists(p, sps, sl, poc, lists)) return rc;
    fill_desc(p, sps, pps, sl, poc, lists, out->desc);
    int slot = -1;  // a slot not held by a reference frame
    for (int k = 0; k < p->dpb_slots && slot < 0; k++) {
        bool used = false;
        for (const Frame &f : p->dpb) used |= (f.short_ref || f.long_ref) && f.slot == k;
        if (!used) slot = k;
    }
    if (slot < 0) return h264g_set_error(H264G_E_ARG, "DPB full: more reference frames than max_num_ref_frames");
    out->desc.dst_slot = slot;
    if (!job.derived) {  // a picture derived ahead never reads the parser's motion field
        if (int rc = start_field(p, sl, lists)) return rc;
    }
    int rc = assemble_slices(p, job, job.ahead_ok, poc, lists, out);
    if (rc == kParseAgain) {
        if ((rc = start_field(p, sl, lists))) return rc;
        rc = assemble_slices(p, job, false, poc, lists, out);
    }
    if (rc) return rc;
    // ---- DPB bookkeeping: the current frame, then marking
    Frame cur;
    cur.slot = out->desc.dst_slot;
    cur.frame_num = sl.frame_num;
    cur.poc = poc;
    if (sl.nal_ref_idc) {
        mark(p, sps, sl, cur);
        if (p->prev_mmco5) {  // 8.2.1: after memory_management_control_operation 5
            cur.frame_num = 0;
            p->prev_ref_top_poc = 0;
            cur.poc = 0;
        }
        cur.mbs = job.mbs;
        p->dpb.erase(std::remove_if(p->dpb.begin(), p->dpb.end(), [](const Frame &f) { return !f.short_ref && !f.long_ref; }),
                     p->dpb.end());
        p->dpb.push_back(cur);
        p->prev_frame_num = cur.frame_num;
    } else {
        p->prev_frame_num = sl.frame_num;
    }
    p->prev_raw_frame_num = sl.frame_num;
    p->last_pic_mmco5 = sl.nal_ref_idc && p->prev_mmco5;
    out->mbs = job.mbs.data();
    out->n_mbs = (uint32_t)job.mbs.size();
    out->levels = job.levels.data();
    out->n_levels = job.levels.size();
    out->width_mbs = W;
    out->height_mbs = H;
    out->dpb_slots = p->dpb_slots;
    out->max_num_reorder_frames = sps.max_num_reorder_frames;
    out->idr = sl.idr;
    out->nal_ref_idc = sl.nal_ref_idc;
    out->frame_num = sl.frame_num;
    out->profile_idc = sps.profile_idc;
    out->constraint_set3 = sps.constraint3;
    out->level_idc = sps.level_idc;
    p->pictures++;
    return 1;
}

}  // namespace

extern "C" int32_t h264g_parser_next(h264g_parser *p, h264g_parsed_picture *out) {
    if (!p || !out) return h264g_set_error(H264G_E_ARG, "NULL argument");
    memset(out, 0, sizeof(*out));
    p->damage[0] = 0;
    p->out_job.reset();
    // keep up to `threads` pictures gathered, the slice data of each parsing on its own thread
    while (!p->gather_end && (int)p->ahead.size() < p->threads) {
        std::unique_ptr<PicJob> job = gather(p);
        if (!job) break;
        if (p->threads > 1 && job->rc == 0) {
            try {
                job->worker = std::thread(parse_ahead, job.get(), p->gathered - 1);
            } catch (const std::exception &) {  // no thread: the picture parses in order
            }
        }
        p->ahead.push_back(std::move(job));
    }
    if (p->ahead.empty()) return 0;
    std::unique_ptr<PicJob> job = std::move(p->ahead.front());
    p->ahead.pop_front();
    if (job->worker.joinable()) job->worker.join();
    if (job->rc) return h264g_set_error(job->rc, "%s", job->err.c_str());
    const int rc = assemble(p, *job, out);
    if (rc == 1) p->out_job = std::move(job);
    return rc;
}
