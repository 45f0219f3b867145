// kernels.hip -- gfx950 kernels of the H.264 picture-reconstruction engine.
//
// One wavefront (64 lanes) reconstructs one macroblock.  Three stages per picture,
// in the reference's order (inter MBs never read the current picture, intra MBs read
// unfiltered neighbours, deblocking runs after the whole picture is built,
// H264PictureBase.cpp:707):
//   inter   -- fully parallel over inter (and I_PCM) MBs: fractional-sample MC from
//              the device-resident DPB + weighted prediction + residual.
//   intra   -- persistent wavefront over intra MBs: each wave takes the next ticket of a
//              queue ordered by t = x + 2y, waits on its A/B/C/D neighbours' progress
//              words, reconstructs, publishes its own word.
//   deblock -- the same wavefront scheme over all MBs with the raster-order edge filter.
// The queues hold MBs of any number of pictures (batching across streams).
#include <hip/hip_runtime.h>

#include <mutex>

#include "h264g_dev.h"

// Device memory seen through global-address-space pointers: loads and stores through
// them are global_* instructions, which count on the vector-memory counter only.  Through
// generic pointers they would be flat_* instructions, which also count on the LDS counter,
// so every LDS wait would wait for them too (a prefetch from HBM would stall the next LDS
// access).
// (The host pass of this file parses the device code too; it sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
#define GAS __attribute__((address_space(1)))
#else
#define GAS
#endif
typedef GAS uint8_t gu8;
template <typename T>
__device__ __forceinline__ GAS T *G(T *p) { return (GAS T *)p; }
template <typename T>
__device__ __forceinline__ const GAS T *G(const T *p) { return (const GAS T *)p; }


namespace {

constexpr int WAVE = 64;

// ------------------------------------------------------------------ tables
// Table 8-13 / 8-14 zig-zag (scan position -> raster index), H264PictureBase.cpp:4542-4744
// (constexpr so that fully unrolled loops index register arrays with constants)
__device__ constexpr uint8_t c_zz4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
// the field scan of Table 8-13 (field MBs of MBAFF frames)
__device__ constexpr uint8_t c_fld4[16] = {0, 4, 1, 8, 12, 5, 9, 13, 2, 6, 10, 14, 3, 7, 11, 15};
__constant__ uint8_t c_zz8[64] = {
    0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14, 21, 28,
    35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
// raster position -> 8x8 zig-zag scan index (inverse of c_zz8)
__constant__ uint8_t c_izz8[64] = {
    0, 1, 5, 6, 14, 15, 27, 28, 2, 4, 7, 13, 16, 26, 29, 42, 3, 8, 12, 17, 25, 30, 41, 43, 9, 11, 18, 24, 31, 40, 44, 53,
    10, 19, 23, 32, 39, 45, 52, 54, 20, 22, 33, 38, 46, 51, 55, 60, 21, 34, 37, 47, 50, 56, 59, 61, 35, 36, 48, 49, 57, 58, 62, 63};
// raster position -> 8x8 field scan index (inverse of Table 8-14's field scan)
__constant__ uint8_t c_ifld8[64] = {
    0, 3, 8, 15, 22, 30, 38, 52, 1, 4, 14, 21, 29, 37, 45, 53, 2, 7, 16, 23, 31, 39, 46, 58, 5, 9, 20, 28, 36, 44, 51, 59,
    6, 13, 24, 32, 40, 47, 54, 60, 10, 17, 25, 33, 41, 48, 55, 61, 11, 18, 26, 34, 42, 49, 56, 62, 12, 19, 27, 35, 43, 50, 57, 63};
// normAdjust values v (8-315, 8-318), H264PictureBase.cpp:4900-4950
__constant__ uint8_t c_v4[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16}, {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
__constant__ uint8_t c_v8[6][6] = {{20, 18, 32, 19, 25, 24}, {22, 19, 35, 21, 28, 26}, {26, 23, 42, 24, 33, 31},
                                   {28, 25, 45, 26, 35, 33}, {32, 28, 51, 30, 40, 38}, {36, 32, 58, 34, 46, 43}};
// Tables 8-16 / 8-17, Deblock.cpp:1330-1409
__constant__ uint8_t c_alpha[52] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 4, 4, 5, 6, 7, 8, 9, 10, 12, 13,
                                    15, 17, 20, 22, 25, 28, 32, 36, 40, 45, 50, 56, 63, 71, 80, 90, 101, 113, 127, 144, 162, 182, 203, 226, 255, 255};
__constant__ uint8_t c_beta[52] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4,
                                   6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15, 15, 16, 16, 17, 17, 18, 18};
__constant__ uint8_t c_tc0[3][52] = {
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 6, 6, 7, 8, 9, 10, 11, 13},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5, 5, 6, 7, 8, 8, 10, 11, 12, 13, 15, 17},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 23, 25}};

__device__ __forceinline__ int clip3(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int clip1(int v) { return clip3(0, 255, v); }
__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }

__device__ __forceinline__ int ls4(int m, int i, int j) {
    int cls = ((i | j) & 1) == 0 ? 0 : (((i & j) & 1) ? 1 : 2);
    return 16 * c_v4[m][cls];
}

// luma4x4BlkIdx -> (x, y) in the MB (6.4.3)
__device__ __forceinline__ int blk_x(int b) { return ((b >> 2) & 1) * 8 + (b & 1) * 4; }
__device__ __forceinline__ int blk_y(int b) { return ((b >> 3) & 1) * 8 + ((b >> 1) & 1) * 4; }

__device__ __forceinline__ int wave_lane() { return threadIdx.x & (WAVE - 1); }

// Orders the LDS accesses of the lanes of one wave (LDS instructions of a wave execute in
// order; the fences keep the compiler from moving memory operations across).  Every
// LDS working set below belongs to one wave, so this is the only barrier the per-MB code
// needs, also when several waves share a workgroup.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------ residuals
// LDS scratch of one wave: the MB's levels as they arrive -- per slot (0 = Intra16x16
// DC, 1 + luma4x4BlkIdx (8x8 block q = slots 1+4q..4+4q), 17/18 = chroma DC Cb/Cr,
// 19 + 4c + b = chroma AC) its 16-bit mask of non-zero scan positions and the index of
// its first value, and the values in stream order -- then the residual planes.
// Luma interpolation of one list (inter MBs, after the residual is built).  Window row rr
// of 4x4 block b holds reference row y0 - 2 + rr at columns x0 - 2 .. x0 + 5, where
// (x0, y0) is the block's integer sample position; b1 is the unrounded horizontal 6-tap
// value between columns x0 + i and x0 + i + 1 (8-241).
struct WaveMc {
    uint32_t raw[16][9][2];   // window columns 0..7, packed bytes
    uint32_t hb[16][5];       // b = Clip1((b1 + 16) >> 5), i = 0..3, of window rows 2..6
    int16_t b1t[16][4][10];   // b1 of column i, window rows 0..8 (+ one pad row): [b][i][rr]
};

struct alignas(16) WaveRes {
    union {
        struct {
            uint32_t sm[28];          // slot: mask (bits 0..15) | index of its first value << 16
            int16_t vals[7 * WAVE];   // level values in stream order (at most 27 * 16)
            int32_t dense[4][64];     // levels: the 8x8 blocks in raster order, or [0] = the
                                      // Intra16x16 DC block in scan order
            int32_t cdc[2][4];        // scan-order levels: chroma DC
            int32_t t8[4][64];        // 8x8 transform: rows after the horizontal pass
        };
        WaveMc mc;                    // live only after the residual planes are built
    };
    int16_t rl[256];      // luma residual, raster 16x16
    int16_t rc[2][64];    // chroma residual, raster 8x8
};

// Slot (see WaveRes) of the MB's i-th level mask, in the order of the level stream: the
// Intra16x16 DC block, the four 4x4 blocks (or 16-level chunks of the 8x8 block) of each
// 8x8 quadrant whose cbp bit is set, the chroma DC blocks, the chroma AC blocks; -1 past
// the end.
__device__ __forceinline__ int level_slot(int i, int i16, int cbp) {
    if (i16) {
        if (i == 0) return 0;
        i--;
    }
    const int lm = cbp & 15, nl = 4 * __popc(lm);
    if (i < nl) {
        int c = i >> 2, q = 0;  // q: the (i >> 2)-th set bit of lm
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int set = (lm >> t) & 1;
            if (set && c == 0) q = t;
            c -= set;
        }
        return 1 + 4 * q + (i & 3);
    }
    i -= nl;
    const int cc = cbp >> 4;
    if (cc >= 1) {
        if (i < 2) return 17 + i;
        i -= 2;
    }
    if (cc == 2 && i < 8) return 19 + i;
    return -1;
}

// Scan positions of the samples (r, 0..3) of a 4x4 block, 4 bits each: the inverse of c_zz4
// (IZZ4) and of c_fld4 (IFLD4).
constexpr uint64_t iscan4_table(const uint8_t (&scan)[16]) {
    uint64_t v = 0;
    for (int k = 0; k < 16; k++) v |= (uint64_t)k << (4 * scan[k]);
    return v;
}
constexpr uint64_t IZZ4 = iscan4_table(c_zz4);
constexpr uint64_t IFLD4 = iscan4_table(c_fld4);

// 4x4 transpose over the four lanes of a quad (lane r holds row r in m[0..3], afterwards
// column r): two exchange stages with DPP quad permutations.
__device__ __forceinline__ int dpp_xor1(int v) { return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true); }  // [1,0,3,2]
__device__ __forceinline__ int dpp_xor2(int v) { return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true); }  // [2,3,0,1]
__device__ __forceinline__ void quad_transpose(int32_t (&m)[4], int r) {
    const bool odd = r & 1, hi = r & 2;
    int a = dpp_xor1(odd ? m[0] : m[1]), b = dpp_xor1(odd ? m[2] : m[3]);
    if (odd) { m[0] = a; m[2] = b; } else { m[1] = a; m[3] = b; }
    a = dpp_xor2(hi ? m[0] : m[2]);
    b = dpp_xor2(hi ? m[1] : m[3]);
    if (hi) { m[0] = a; m[1] = b; } else { m[2] = a; m[3] = b; }
}

// 4x4 byte transpose over the four lanes of a quad: lane j holds row j of a block of bytes
// (byte c = column c) and gets column j (byte r = row r).  Two exchange stages, each a DPP
// quad permutation and one byte permute with a lane-dependent selector; applied twice it
// gives the rows back.
__device__ __forceinline__ uint32_t quad_transpose_bytes(uint32_t v, int j) {
    // even j: (own c0, partner c0, own c2, partner c2); odd j: (partner c1, own c1, partner c3, own c3)
    const uint32_t x = (uint32_t)dpp_xor1((int)v);
    const uint32_t t = __builtin_amdgcn_perm(x, v, (j & 1) ? 0x03070105u : 0x06020400u);
    // j < 2: own low half, then the low half of lane j + 2; j >= 2: the high half of lane j - 2, then own
    const uint32_t y = (uint32_t)dpp_xor2((int)t);
    return __builtin_amdgcn_perm(y, t, (j & 2) ? 0x03020706u : 0x05040100u);
}

// Level at scan position k (< 16) of the block with slot entry e: 0 unless mask bit k is
// set, else the value at the block's first index plus the number of mask bits below k.
__device__ __forceinline__ int32_t level_at(uint32_t e, const int16_t *vals, int k) {
    const uint32_t below = __popc(e & ((1u << k) - 1u));
    return ((e >> k) & 1u) ? (int32_t)vals[(e >> 16) + below] : 0;
}

// 8.5.12 for row r of a 4x4 block on the quad that holds the block (lane r of the quad):
// dequantise the row's levels (slot entry e), horizontal pass, transpose, vertical pass;
// returns column r of the residual in out[0..3] (rows 0..3).  dc_pass: sample (0, 0) is
// `dc` as it stands (Intra16x16 and chroma DC).  field: the levels follow the field scan
// (field MBs, H264PictureBase.cpp:3419).
template <bool SCALED = false>
__device__ __forceinline__ void residual4x4_quad(uint32_t se, const int16_t *vals, int qp, bool dc_pass, int dc, int r,
                                                 bool field, int32_t (&out)[4], const GAS uint16_t *lsm = nullptr) {
    const int m = qp % 6, e = qp / 6;
    // LevelScale4x4 of row r: with flat lists 16 v0 at (even, even), 16 v1 at (odd, odd),
    // 16 v2 elsewhere; with scaling matrices (SCALED) lsm holds LevelScale4x4(qP % 6, i, j),
    // raster
    int ls[4];
    if (SCALED) {
        const uint2 t = *(const GAS uint2 *)(lsm + 4 * r);
        ls[0] = (int)(t.x & 0xFFFFu); ls[1] = (int)(t.x >> 16); ls[2] = (int)(t.y & 0xFFFFu); ls[3] = (int)(t.y >> 16);
    } else {
        const int s0 = 16 * c_v4[m][0], s1 = 16 * c_v4[m][1], s2 = 16 * c_v4[m][2];
        ls[0] = ls[2] = (r & 1) ? s2 : s0;
        ls[1] = ls[3] = (r & 1) ? s1 : s2;
    }
    // 8-336 / 8-337 as one expression: qP / 6 >= 4 shifts left, else rounds and shifts right
    const int lsh = e >= 4 ? e - 4 : 0, rsh = e >= 4 ? 0 : 4 - e, rnd = e >= 4 ? 0 : 1 << (3 - e);
    const uint32_t iz = (uint32_t)((field ? IFLD4 : IZZ4) >> (16 * r)) & 0xFFFFu;
    int32_t d[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int32_t v = level_at(se, vals, (int)((iz >> (4 * j)) & 15u));
        d[j] = (int32_t)(((uint32_t)(v * ls[j]) << lsh) + (uint32_t)rnd) >> rsh;
    }
    if (dc_pass && r == 0) d[0] = dc;
    int32_t f[4];
    {
        const int32_t e0 = d[0] + d[2], e1 = d[0] - d[2], e2 = (d[1] >> 1) - d[3], e3 = d[1] + (d[3] >> 1);
        f[0] = e0 + e3; f[1] = e1 + e2; f[2] = e1 - e2; f[3] = e0 - e3;
    }
    quad_transpose(f, r);
    const int32_t g0 = f[0] + f[2], g1 = f[0] - f[2], g2 = (f[1] >> 1) - f[3], g3 = f[1] + (f[3] >> 1);
    out[0] = (g0 + g3 + 32) >> 6; out[1] = (g1 + g2 + 32) >> 6;
    out[2] = (g1 - g2 + 32) >> 6; out[3] = (g0 - g3 + 32) >> 6;
}

__device__ __forceinline__ void idct8(int32_t *v, int s) {  // 8.5.13.2, in place, stride s
    int32_t d0 = v[0], d1 = v[s], d2 = v[2 * s], d3 = v[3 * s], d4 = v[4 * s], d5 = v[5 * s], d6 = v[6 * s], d7 = v[7 * s];
    int32_t e0 = d0 + d4, e1 = -d3 + d5 - d7 - (d7 >> 1), e2 = d0 - d4, e3 = d1 + d7 - d3 - (d3 >> 1);
    int32_t e4 = (d2 >> 1) - d6, e5 = -d1 + d7 + d5 + (d5 >> 1), e6 = d2 + (d6 >> 1), e7 = d3 + d5 + d1 + (d1 >> 1);
    int32_t f0 = e0 + e6, f1 = e1 + (e7 >> 2), f2 = e2 + e4, f3 = e3 + (e5 >> 2);
    int32_t f4 = e2 - e4, f5 = (e3 >> 2) - e5, f6 = e0 - e6, f7 = e7 - (e1 >> 2);
    v[0] = f0 + f7; v[s] = f2 + f5; v[2 * s] = f4 + f3; v[3 * s] = f6 + f1;
    v[4 * s] = f6 - f1; v[5 * s] = f4 - f3; v[6 * s] = f2 - f5; v[7 * s] = f0 - f7;
}

// Places the MB's levels in LDS and computes all residual blocks into ws->rl / ws->rc.  Must be called by the whole wave.  Returns false when the MB has no levels.
// Field MBs of MBAFF frames take the field scans.  lstab: the picture's LevelScale tables
// when it has scaling matrices (DevPic::ls), else null (flat lists).
template <bool MBAFF, bool SCALED>
__device__ __forceinline__ bool build_residual(const h264g_mb *mb, const DevAux &ax, const int16_t *levels, WaveRes *ws,
                                               const GAS uint16_t *lstab) {
    const int lane = wave_lane();
    if (!(ax.kind & AUX_RESID)) return false;
    const bool field = MBAFF && (mb->flags & H264G_MBF_FIELD) != 0;
    const int cbp = mb->cbp;
    const int i16 = mb->mb_type == H264G_MB_I16x16;
    const int t8 = (ax.kind & AUX_T8) ? 1 : 0;
    const int qp = mb->qp_y;
    // the MB's luma lists (intra or inter), which serve its chroma blocks too (quirk Q13)
    const bool inter = !(ax.kind & AUX_INTRA);
    const GAS uint16_t *ls4t = SCALED ? lstab + (inter ? LS4_LIST : 0) : nullptr;
    const GAS uint16_t *ls8t = SCALED ? lstab + LS8_AT + (inter ? LS8_LIST : 0) : nullptr;
    const int16_t *lv = levels + mb->level_offset;
    // <= 27 masks of 16 bits, so at most 432 values; the host counted them, so the masks
    // and the values load together (coalesced, all in flight at once)
    const int nmask = ax.nmask, total = (int)(ax.ires >> AUX_NVALS_SHIFT);
    const int mask = lane < nmask ? (uint16_t)lv[lane] : 0;
    int16_t v[7];
#pragma unroll
    for (int t = 0; t < 7; t++) {
        const int k = lane + t * WAVE;
        v[t] = k < total ? lv[nmask + k] : (int16_t)0;
    }
    const int cnt = __popc(mask);
    // exclusive prefix sum of cnt (0..16) over the lanes, bit plane by bit plane (ballot +
    // mbcnt, no cross-lane shuffles)
    int pre = 0;
#pragma unroll
    for (int bit = 0; bit < 5; bit++) {
        const uint64_t bl = __builtin_amdgcn_ballot_w64((cnt >> bit) & 1);
        pre += (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bl >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bl, 0u)) << bit;
    }
    // slot entries (absent slots 0; the wave's LDS writes land in program order) and values
    if (lane < 28) ws->sm[lane] = 0u;
    if (lane < nmask) ws->sm[level_slot(lane, i16, cbp)] = (uint32_t)mask | (uint32_t)pre << 16;
#pragma unroll
    for (int t = 0; t < 7; t++)
        if (lane + t * WAVE < total) ws->vals[lane + t * WAVE] = v[t];
    wave_sync();
    // dense levels where a pass reads positions of other rows: the 8x8 blocks, the
    // Intra16x16 DC block (lanes 0..15) and the chroma DC blocks (lanes 56..63)
    if (t8) {  // lane = raster position in the 8x8 blocks, p its scan position
        const int p = field ? c_ifld8[lane] : c_izz8[lane];
#pragma unroll
        for (int q = 0; q < 4; q++) ws->dense[q][lane] = level_at(ws->sm[1 + 4 * q + (p >> 4)], ws->vals, p & 15);
    } else if (i16 && lane < 16) {
        ws->dense[0][lane] = level_at(ws->sm[0], ws->vals, lane);
    }
    if ((cbp >> 4) && lane >= 56) ws->cdc[(lane - 56) >> 2][lane & 3] = level_at(ws->sm[17 + ((lane - 56) >> 2)], ws->vals, lane & 3);
    if (t8 || i16 || (cbp >> 4)) wave_sync();
    if (t8) {
        // 8.5.13: lanes 0-31 = (8x8 block q, row i): dequantise + horizontal pass ...
        if (lane < 32) {
            const int q = lane >> 3, i = lane & 7, m = qp % 6, e = qp / 6;
            const int32_t *src = &ws->dense[q][i * 8];
            // LevelScale8x8 of row i by column class (j % 4 == 0, j odd, j % 4 == 2): the
            // six normAdjust8x8 classes of 8-318
            const int v0 = 16 * c_v8[m][0], v1 = 16 * c_v8[m][1], v2 = 16 * c_v8[m][2], v3 = 16 * c_v8[m][3],
                      v4 = 16 * c_v8[m][4], v5 = 16 * c_v8[m][5];
            const int ls_0 = (i & 1) ? v3 : (i & 2) ? v4 : v0;
            const int ls_1 = (i & 1) ? v1 : (i & 2) ? v5 : v3;
            const int ls_2 = (i & 1) ? v5 : (i & 2) ? v2 : v4;
            int lsr[8];  // LevelScale8x8 of row i: the classes above, or the matrix row
#pragma unroll
            for (int j = 0; j < 8; j++) lsr[j] = (j & 1) ? ls_1 : (j & 2) ? ls_2 : ls_0;
            if (SCALED) {
                const uint4 t = *(const GAS uint4 *)(ls8t + 64 * m + 8 * i);
                const uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
                for (int j = 0; j < 8; j++) lsr[j] = (int)((w[j >> 1] >> (16 * (j & 1))) & 0xFFFFu);
            }
            // 8-338 / 8-339 as one expression (qP / 6 >= 6 shifts left)
            const int lsh = e >= 6 ? e - 6 : 0, rsh = e >= 6 ? 0 : 6 - e, rnd = e >= 6 ? 0 : 1 << (5 - e);
            int32_t d[8];
#pragma unroll
            for (int j = 0; j < 8; j++) d[j] = (int32_t)(((uint32_t)(src[j] * lsr[j]) << lsh) + (uint32_t)rnd) >> rsh;
            idct8(d, 1);
#pragma unroll
            for (int j = 0; j < 8; j++) ws->t8[q][i * 8 + j] = d[j];
        }
    } else if (i16 || (cbp & 15)) {
        // 4x4 luma blocks: the quad of lanes 4b..4b+3 takes block luma4x4BlkIdx b, lane 4b + r
        // its row r (then column r)
        const int b = lane >> 2, r = lane & 3;
        int dc = 0;
        if (i16) {  // 8.5.10: f = (H c H)[row][col] with H the 4x4 Hadamard, then scaling
            const int row = blk_y(b) >> 2, col = blk_x(b) >> 2;
            // H[r][i] is -1 where: r = 1, i >= 2; r = 2, i in {1, 2}; r = 3, i odd
            auto hneg = [](int r, int i) { return r == 1 ? (i >> 1) : r == 2 ? ((i == 1) | (i == 2)) : r == 3 ? (i & 1) : 0; };
            int32_t f = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                int32_t acc = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    int32_t vz = 0, vf = 0;
#pragma unroll
                    for (int k = 0; k < 16; k++) {  // resolved at compile time
                        if (c_zz4[k] == i * 4 + j) vz = ws->dense[0][k];
                        if (c_fld4[k] == i * 4 + j) vf = ws->dense[0][k];
                    }
                    const int32_t v = field ? vf : vz;
                    acc += hneg(col, j) ? -v : v;
                }
                f += hneg(row, i) ? -acc : acc;
            }
            const int sc = SCALED ? ls4t[16 * (qp % 6)] : ls4(qp % 6, 0, 0);
            dc = qp >= 36 ? (f * sc) << (qp / 6 - 6) : (f * sc + (1 << (5 - qp / 6))) >> (6 - qp / 6);
        }
        int32_t o[4];
        residual4x4_quad<SCALED>(ws->sm[1 + b], ws->vals, qp, i16, dc, r, field, o, ls4t + 16 * (qp % 6));
        const int x0 = blk_x(b), y0 = blk_y(b);
#pragma unroll
        for (int i = 0; i < 4; i++) ws->rl[(y0 + i) * 16 + x0 + r] = (int16_t)o[i];
    } else {  // no luma levels
        ((uint2 *)ws->rl)[lane] = make_uint2(0u, 0u);
    }
    if (!(cbp >> 4)) {  // no chroma levels
        ((uint32_t *)ws->rc)[lane] = 0u;
    } else if (lane < 32) {  // chroma: 8.5.11 DC then 8.5.12 per 4x4 block, a quad per block
        const int c = lane >> 4, b = (lane >> 2) & 3, r = lane & 3;
        const int qpc = ax.qpc[c];
        const int32_t *dcv = ws->cdc[c];
        const int32_t e00 = dcv[0] + dcv[2], e01 = dcv[1] + dcv[3], e10 = dcv[0] - dcv[2], e11 = dcv[1] - dcv[3];
        const int32_t f = b == 0 ? e00 + e01 : b == 1 ? e00 - e01 : b == 2 ? e10 + e11 : e10 - e11;
        const int dc = ((f * (SCALED ? ls4t[16 * (qpc % 6)] : ls4(qpc % 6, 0, 0))) << (qpc / 6)) >> 5;
        int32_t o[4];
        residual4x4_quad<SCALED>(ws->sm[19 + 4 * c + b], ws->vals, qpc, true, dc, r, field, o, ls4t + 16 * (qpc % 6));
        const int x0 = (b & 1) * 4, y0 = (b >> 1) * 4;
#pragma unroll
        for (int i = 0; i < 4; i++) ws->rc[c][(y0 + i) * 8 + x0 + r] = (int16_t)o[i];
    }
    if (t8) {
        wave_sync();
        // ... then lanes 0-31 = (block q, column j): vertical pass and (x + 32) >> 6
        if (lane < 32) {
            const int q = lane >> 3, j = lane & 7;
            int32_t d[8];
#pragma unroll
            for (int i = 0; i < 8; i++) d[i] = ws->t8[q][i * 8 + j];
            idct8(d, 1);
            const int x0 = (q & 1) * 8, y0 = (q >> 1) * 8;
#pragma unroll
            for (int i = 0; i < 8; i++) ws->rl[(y0 + i) * 16 + x0 + j] = (int16_t)((d[i] + 32) >> 6);
        }
    }
    wave_sync();
    return true;
}

// ------------------------------------------------------------ inter MC
__device__ __forceinline__ gu8 *slot_base(const DevPic &P, int s) { return G(P.dpb) + (uint64_t)s * P.slot_bytes; }

// Where an MB's samples lie (6.4.1, H264PictureBase.cpp:2503-2530): the picture row of its
// first luma / chroma sample and the row step (2 for field MBs of MBAFF frames, whose rows
// are every other picture row, 8.5.14), and its row origin in the coordinates of the
// pictures it predicts from (frame rows, or field rows for field MBs,
// H264InterPrediction.cpp:553-557).
struct MbGeo {
    int mbx, ly, cy, step, mcy, bot;
    bool field;
};
template <bool MBAFF>
__device__ __forceinline__ MbGeo mb_geo(const DevPic &P, int addr) {
    MbGeo g;
    if (!MBAFF || !P.mbaff) {
        const int y = addr / P.W;
        g.mbx = addr % P.W;
        g.ly = y * 16;
        g.cy = y * 8;
        g.step = 1;
        g.mcy = g.ly;
        g.bot = 0;
        g.field = false;
        return g;
    }
    const int pair = addr >> 1, py = pair / P.W;
    g.mbx = pair % P.W;
    g.bot = addr & 1;
    g.field = (G(P.mbs)[addr].flags & H264G_MBF_FIELD) != 0;
    g.step = g.field ? 2 : 1;
    g.ly = py * 32 + (g.field ? g.bot : 16 * g.bot);
    g.cy = py * 16 + (g.field ? g.bot : 8 * g.bot);
    g.mcy = g.field ? py * 16 : g.ly;
    return g;
}

// (P1 + P2 + 1) >> 1 of every luma position (xF, yF), 8.4.2.2.1 (H264InterPrediction.cpp:2344-2480),
// with P1, P2 among 0 = G (integer sample), 1 = b (horizontal half sample), 2 = h
// (vertical half sample), 3 = j: 4 bits (P1 | P2 << 2) at bit 4 * (4 * yF + xF).  The
// quarter positions right of or below the half samples read them one column right
// (m instead of h, G at x + 1) or one row down (s instead of b, G at y + 1).
constexpr uint64_t mc_sel_table() {
    const int t[16][2] = {{0, 0}, {0, 1}, {1, 1}, {0, 1}, {0, 2}, {1, 2}, {1, 3}, {1, 2},
                          {2, 2}, {2, 3}, {3, 3}, {2, 3}, {0, 2}, {1, 2}, {1, 3}, {1, 2}};
    uint64_t v = 0;
    for (int i = 0; i < 16; i++) v |= (uint64_t)(t[i][0] | t[i][1] << 2) << (4 * i);
    return v;
}
constexpr uint64_t MC_SEL = mc_sel_table();
// 6-tap (1, -5, 20, 20, -5, 1) as signed byte weights for v_dot4: samples s..s+3 and
// s+4, s+5 (bytes 0-1; bytes 2-3 for the second column of a pair).  The samples enter
// as x ^ 0x80 = x - 128, so every tap sum is short by 128 * 32 = 4096.
constexpr int TAP_LO = 0x1414fb01, TAP_HI = 0x000001fb, TAP_HI2 = 0x01fb0000, TAP_BIAS = 4096;

__device__ __forceinline__ uint32_t align8(uint32_t hi, uint32_t lo, uint32_t bytes) {
    return __builtin_amdgcn_alignbyte(hi, lo, bytes);
}
__device__ __forceinline__ int sdot4(uint32_t a, int w, int acc) { return __builtin_amdgcn_sdot4((int)a, w, acc, false); }
__device__ __forceinline__ int sdot2(uint32_t a, short w0, short w1, int acc) {
    typedef short short2_t __attribute__((ext_vector_type(2)));
    return __builtin_amdgcn_sdot2(__builtin_bit_cast(short2_t, a), (short2_t){w0, w1}, acc, false);
}
// Low bytes of v0..v3 as one dword, through v_perm.  Written with shifts and ORs, hipcc
// fuses Clip1(x >> n) of two values into v_ashr_pk_u8_i32 and ORs the other two bytes
// onto its destination, whose upper half then still holds the register's old contents
// on gfx950; the permutes keep that instruction out.
__device__ __forceinline__ uint32_t pack4(int v0, int v1, int v2, int v3) {
    const uint32_t lo = __builtin_amdgcn_perm((uint32_t)v1, (uint32_t)v0, 0x0c0c0400u);
    const uint32_t hi = __builtin_amdgcn_perm((uint32_t)v3, (uint32_t)v2, 0x0c0c0400u);
    return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}
// Per-byte (a + b + 1) >> 1.
__device__ __forceinline__ uint32_t avg4(uint32_t a, uint32_t b) { return (a | b) - (((a ^ b) >> 1) & 0x7f7f7f7fu); }

// Luma prediction samples of list l: lane = (4x4 block b = lane & 15, row r = lane >> 4),
// four samples packed in a dword (0 where the lane's quadrant does not use list l).
// Stage 1 (lane = block b, window rows r, r + 4, r + 8): dword loads of the window row,
// the horizontal 6-tap values b1 of its four columns (v_dot4), all into LDS.  Stage 2:
// the lane's vertical 6-taps on window bytes (h, or m) and on b1 (j), and the pair of
// the position's inputs.  Both stages are the same for every (xF, yF), so lanes with
// different quarter-sample positions do not diverge; the wave skips what no lane uses.
// A field MB of an MBAFF frame predicts from one field of the reference frame: its rows
// are every other frame row from row `parity`, and the reference's field view reports
// that row pitch, twice the width, as its width, which is also where it clamps x
// (H264InterPrediction.cpp:2363 with H264PictureBase.cpp:209-210): right of the picture
// the samples come from the next frame row, on the bottom field's last row from the next
// plane (the DPB slots of MBAFF streams end in a zeroed tail for the Cr plane).
__device__ __forceinline__ uint32_t luma_list(const DevPic &P, const h264g_mb *mb, const DevAux &ax, int l, const MbGeo &g,
                                              WaveRes *ws) {
    const int lane = wave_lane();
    const int W = P.W, wl = W * 16 * g.step, hl = P.H * 16 / g.step;  // row pitch / clamp width, clamp height
    const int bx0 = g.mbx * 16, by0 = g.mcy;
    const int b = lane & 15, r = lane >> 4, bx = b & 3, by = b >> 2, q = ((by >> 1) << 1) | (bx >> 1);
    const bool use = (ax.pf >> (4 * l + q)) & 1;
    const int mvx = mb->mv[l][b][0], mvy = mb->mv[l][b][1];
    const int xF = mvx & 3, yF = mvy & 3;
    const int sel = (int)((MC_SEL >> (4 * (4 * yF + xF))) & 15), s1 = sel & 3, s2 = sel >> 2;
    const uint32_t uses = use ? (1u << s1) | (1u << s2) : 0u;
    const bool need_b = __ballot(uses & 2) != 0, need_h = __ballot(uses & 4) != 0, need_j = __ballot(uses & 8) != 0;
    const int x0 = bx0 + bx * 4 + (mvx >> 2), y0 = by0 + by * 4 + (mvy >> 2);
    WaveMc &mc = ws->mc;
    if (use) {
        const uint8_t *ref = slot_base(P, ax.refid[l][q]) + (g.field && ((ax.fpar >> (4 * l + q)) & 1) ? wl / 2 : 0);
        const int xs = x0 - 2, xa = xs & ~3, off = xs & 3;
        const bool inside = xa >= 0 && xa + 12 <= wl;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int rr = r + 4 * k;
            if (rr < 9) {
                const uint8_t *row = ref + (uint64_t)clip3(0, hl - 1, y0 - 2 + rr) * wl;
                uint32_t w0, w1, w2;
                if (inside) {
                    const uint32_t *p = (const uint32_t *)(row + xa);
                    const uint32_t d0 = p[0], d1 = p[1], d2 = p[2];
                    w0 = align8(d1, d0, off); w1 = align8(d2, d1, off); w2 = d2 >> (8 * off);
                } else {  // the window leaves the picture: clamped columns (8.4.2.2.1)
                    int v[9];
#pragma unroll
                    for (int t = 0; t < 9; t++) v[t] = row[clip3(0, wl - 1, xs + t)];
                    w0 = pack4(v[0], v[1], v[2], v[3]); w1 = pack4(v[4], v[5], v[6], v[7]); w2 = (uint32_t)v[8];
                }
                mc.raw[b][rr][0] = w0;
                mc.raw[b][rr][1] = w1;
                if (need_b || need_j) {
                    const uint32_t s0 = w0 ^ 0x80808080u, sx = w1 ^ 0x80808080u, sy = w2 ^ 0x80u;
                    int hv[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const uint32_t lo = i ? align8(sx, s0, i) : s0, hi = i ? align8(sy, sx, i) : sx;
                        hv[i] = sdot4(hi, TAP_HI, sdot4(lo, TAP_LO, TAP_BIAS));
                        mc.b1t[b][i][rr] = (int16_t)hv[i];
                    }
                    if (rr >= 2 && rr <= 6)
                        mc.hb[b][rr - 2] = pack4(clip1((hv[0] + 16) >> 5), clip1((hv[1] + 16) >> 5), clip1((hv[2] + 16) >> 5),
                                                 clip1((hv[3] + 16) >> 5));
                }
            }
        }
    }
    wave_sync();
    uint32_t out = 0;
    if (use) {
        const int dx = xF == 3, dy = yF == 3;
        uint32_t R[6];
#pragma unroll
        for (int k = 0; k < 6; k++) R[k] = align8(mc.raw[b][r + k][1], mc.raw[b][r + k][0], 2 + dx);
        uint32_t cand[4];
        cand[0] = dy ? R[3] : R[2];
        cand[1] = need_b ? mc.hb[b][r + dy] : 0u;
        cand[2] = 0;
        if (need_h) {  // vertical 6-tap of the four columns: byte transpose, then v_dot4
            const uint32_t t0 = __builtin_amdgcn_perm(R[1], R[0], 0x05010400u), t1 = __builtin_amdgcn_perm(R[1], R[0], 0x07030602u);
            const uint32_t t2 = __builtin_amdgcn_perm(R[3], R[2], 0x05010400u), t3 = __builtin_amdgcn_perm(R[3], R[2], 0x07030602u);
            const uint32_t c0 = __builtin_amdgcn_perm(t2, t0, 0x05040100u) ^ 0x80808080u;
            const uint32_t c1 = __builtin_amdgcn_perm(t2, t0, 0x07060302u) ^ 0x80808080u;
            const uint32_t c2 = __builtin_amdgcn_perm(t3, t1, 0x05040100u) ^ 0x80808080u;
            const uint32_t c3 = __builtin_amdgcn_perm(t3, t1, 0x07060302u) ^ 0x80808080u;
            const uint32_t e01 = __builtin_amdgcn_perm(R[5], R[4], 0x05010400u) ^ 0x80808080u;
            const uint32_t e23 = __builtin_amdgcn_perm(R[5], R[4], 0x07030602u) ^ 0x80808080u;
            const int h0 = sdot4(e01, TAP_HI, sdot4(c0, TAP_LO, TAP_BIAS + 16));
            const int h1 = sdot4(e01, TAP_HI2, sdot4(c1, TAP_LO, TAP_BIAS + 16));
            const int h2 = sdot4(e23, TAP_HI, sdot4(c2, TAP_LO, TAP_BIAS + 16));
            const int h3 = sdot4(e23, TAP_HI2, sdot4(c3, TAP_LO, TAP_BIAS + 16));
            cand[2] = pack4(clip1(h0 >> 5), clip1(h1 >> 5), clip1(h2 >> 5), clip1(h3 >> 5));
        }
        cand[3] = 0;
        if (need_j) {  // vertical 6-tap over b1 (8-243) with v_dot2 on row pairs
            const uint32_t sh = 2 * (r & 1);
            int jv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t *c = (const uint32_t *)&mc.b1t[b][i][r & ~1];
                const uint32_t d0 = c[0], d1 = c[1], d2 = c[2], d3 = c[3];
                const uint32_t p0 = align8(d1, d0, sh), p1 = align8(d2, d1, sh), p2 = align8(d3, d2, sh);
                jv[i] = clip1(sdot2(p2, -5, 1, sdot2(p1, 20, 20, sdot2(p0, 1, -5, 512))) >> 10);
            }
            cand[3] = pack4(jv[0], jv[1], jv[2], jv[3]);
        }
        const uint32_t a = s1 == 0 ? cand[0] : s1 == 1 ? cand[1] : s1 == 2 ? cand[2] : cand[3];
        const uint32_t c = s2 == 0 ? cand[0] : s2 == 1 ? cand[1] : s2 == 2 ? cand[2] : cand[3];
        out = avg4(a, c);
    }
    wave_sync();  // the buffer is reused by the next list
    return out;
}

// Two horizontally adjacent chroma prediction samples at (xC, yC) with a 1/8-sample
// MV, 8.4.2.2.2 (H264InterPrediction.cpp:2485-2522), coordinates clamped: samples
// A, B (row yi) and C, D (row yi + 1) weighted (8 - xF)(8 - yF), xF(8 - yF), (8 - xF)yF,
// xF yF.  The three columns of each row arrive as one dword (two aligned dword loads
// where the columns lie inside the row, clamped byte loads at the picture's edges), and
// each output is one v_dot4 of the bytes (A, B, C, D) with the four weights.
__device__ __forceinline__ void chroma2(const gu8 *ref, int wc, int hc, int xC, int yC, int mvx, int mvy, int &o0, int &o1) {
    const int xi = xC + (mvx >> 3), yi = yC + (mvy >> 3), xF = mvx & 7, yF = mvy & 7;
    const gu8 *r0 = ref + (uint64_t)clip3(0, hc - 1, yi) * wc, *r1 = ref + (uint64_t)clip3(0, hc - 1, yi + 1) * wc;
    const int xa = xi & ~3;
    uint32_t a, b;  // bytes 0..2: columns xi .. xi + 2 of rows yi, yi + 1
    if (xa >= 0 && xa + 8 <= wc) {
        const GAS uint32_t *p0 = (const GAS uint32_t *)(r0 + xa), *p1 = (const GAS uint32_t *)(r1 + xa);
        a = align8(p0[1], p0[0], (uint32_t)(xi & 3));
        b = align8(p1[1], p1[0], (uint32_t)(xi & 3));
    } else {
        const int x0 = clip3(0, wc - 1, xi), x1 = clip3(0, wc - 1, xi + 1), x2 = clip3(0, wc - 1, xi + 2);
        a = pack4(r0[x0], r0[x1], r0[x2], 0);
        b = pack4(r1[x0], r1[x1], r1[x2], 0);
    }
    const uint32_t w = (uint32_t)((8 - xF) * (8 - yF)) | (uint32_t)(xF * (8 - yF)) << 8 | (uint32_t)((8 - xF) * yF) << 16 |
                       (uint32_t)(xF * yF) << 24;
    o0 = (int)(__builtin_amdgcn_udot4(__builtin_amdgcn_perm(b, a, 0x05040100u), w, 32u, false) >> 6);
    o1 = (int)(__builtin_amdgcn_udot4(__builtin_amdgcn_perm(b, a, 0x06050201u), w, 32u, false) >> 6);
}

// Explicit weighted prediction, 8.4.2.3.2 (H264InterPrediction.cpp:2526-2829), as one
// v_dot2 per sample: the two lists' samples as the 16-bit halves of a dword (list 0
// low), weighted (w0, w1) -- one of them 0 for a single list -- plus the rounding term,
// then shifted, offset and clipped.  The host resolved the weights, including the
// L1-only luma entries of quirk Q2, into w0 / o0.  (Default weighting, 8.4.2.3.1, is
// the rounded average and stays outside.)
struct WpTerms {
    uint32_t w;          // (w0, w1) as 16-bit halves
    int rnd, sh, off;
};
__device__ __forceinline__ WpTerms wp_terms(const DevWeight &w, int c, int pf0, int pf1) {
    const int lw = w.logWD[c];
    WpTerms t;
    if (pf0 && pf1) {
        t.w = (uint16_t)w.w0[c] | (uint32_t)(uint16_t)w.w1[c] << 16;
        t.rnd = 1 << lw;
        t.sh = lw + 1;
        t.off = (w.o0[c] + w.o1[c] + 1) >> 1;
    } else {
        const int ww = (pf0 || c == 0) ? w.w0[c] : w.w1[c];
        t.w = pf0 ? (uint32_t)(uint16_t)ww : (uint32_t)(uint16_t)ww << 16;
        t.rnd = lw >= 1 ? 1 << (lw - 1) : 0;
        t.sh = lw;
        t.off = (pf0 || c == 0) ? w.o0[c] : w.o1[c];
    }
    return t;
}
__device__ __forceinline__ int wp_apply(const WpTerms &t, uint32_t pair) {
    typedef short short2_t __attribute__((ext_vector_type(2)));
    const int v = __builtin_amdgcn_sdot2(__builtin_bit_cast(short2_t, pair), __builtin_bit_cast(short2_t, t.w), t.rnd, false);
    return clip1((v >> t.sh) + t.off);
}

// Inter MB (or I_PCM copy): prediction + residual, written to the destination slot.
// SCALED: the launch holds pictures with scaling matrices (their residual code reads the
// LevelScale tables; the other pictures' the flat one).
template <bool MBAFF, bool SCALED>
__device__ void inter_mb(const DevPic &P, int addr, WaveRes *ws) {
    const int lane = wave_lane();
    const h264g_mb *mb = &G(P.mbs)[addr];
    const DevAux &ax = G(P.aux)[addr];  // per-lane fields (refid[q], wp[q]) are read from memory
    const int W = P.W, H = P.H, wl = W * 16, hl = H * 16, wc = W * 8, hc = H * 8;
    const MbGeo g = mb_geo<MBAFF>(P, addr);
    const int mbx = g.mbx;
    uint8_t *dst = slot_base(P, P.dst);
    uint8_t *dy = dst, *dcb = dst + (uint64_t)wl * hl, *dcr = dcb + (uint64_t)wc * hc;
    if (ax.kind & AUX_PCM) {  // 8.3.5 (H264PictureBase.cpp:2449-2499)
        const int16_t *pcm = G(P.levels) + mb->level_offset;
        for (int i = lane; i < 256; i += WAVE) dy[(g.ly + (i >> 4) * g.step) * wl + mbx * 16 + (i & 15)] = (uint8_t)pcm[i];
        dcb[(g.cy + (lane >> 3) * g.step) * wc + mbx * 8 + (lane & 7)] = (uint8_t)pcm[256 + lane];
        dcr[(g.cy + (lane >> 3) * g.step) * wc + mbx * 8 + (lane & 7)] = (uint8_t)pcm[320 + lane];
        return;
    }
    const bool res = SCALED && P.ls ? build_residual<MBAFF, SCALED>(mb, ax, G(P.levels), ws, G(P.ls))
                                    : build_residual<MBAFF, false>(mb, ax, G(P.levels), ws, nullptr);
    // ---- luma: lane -> 4x4 block b = lane & 15 (raster), row r = lane >> 4
    {
        const int b = lane & 15, r = lane >> 4, bx = b & 3, by = b >> 2, q = (by >> 1) * 2 + (bx >> 1);
        const int pf0 = (ax.pf >> q) & 1, pf1 = (ax.pf >> (4 + q)) & 1;
        uint32_t p0 = 0, p1 = 0;
#pragma nounroll
        for (int l = 0; l < 2; l++) {  // one inlined copy of the interpolation for both lists
            if (!(ax.pf & (0x0F << (4 * l)))) continue;  // MB-uniform
            const uint32_t v = luma_list(P, mb, ax, l, g, ws);
            if (l) p1 = v; else p0 = v;
        }
        const DevWeight &w = G(P.wt)[ax.wp[q]];
        uint32_t packed;
        if (!w.mode) {
            packed = pf0 && pf1 ? avg4(p0, p1) : (pf0 ? p0 : p1);
        } else {  // byte i of p0 / p1 as the 16-bit halves of one dword (v_perm, 0x0c = zero)
            const WpTerms t = wp_terms(w, 0, pf0, pf1);
            packed = pack4(wp_apply(t, __builtin_amdgcn_perm(p1, p0, 0x0c040c00u)),
                           wp_apply(t, __builtin_amdgcn_perm(p1, p0, 0x0c050c01u)),
                           wp_apply(t, __builtin_amdgcn_perm(p1, p0, 0x0c060c02u)),
                           wp_apply(t, __builtin_amdgcn_perm(p1, p0, 0x0c070c03u)));
        }
        if (res) {
            const int16_t *rl = &ws->rl[(by * 4 + r) * 16 + bx * 4];
            packed = pack4(clip1((int)(packed & 255) + rl[0]), clip1((int)((packed >> 8) & 255) + rl[1]),
                           clip1((int)((packed >> 16) & 255) + rl[2]), clip1((int)(packed >> 24) + rl[3]));
        }
        *(uint32_t *)&dy[(uint64_t)(g.ly + (by * 4 + r) * g.step) * wl + mbx * 16 + bx * 4] = packed;
    }
    // ---- chroma: lane -> component c, 4x4 luma block b, chroma row r (2 samples)
    {
        int c = lane >> 5, i = lane & 31, b = i >> 1, r = i & 1, bx = b & 3, by = b >> 2, q = (by >> 1) * 2 + (bx >> 1);
        int pf0 = (ax.pf >> q) & 1, pf1 = (ax.pf >> (4 + q)) & 1;
        int xC = mbx * 8 + bx * 2, yC = g.mcy / 2 + by * 2 + r;
        int c00 = 0, c01 = 0, c10 = 0, c11 = 0;
        const uint64_t coff = (uint64_t)wl * hl + (uint64_t)c * wc * hc;
        // field MBs: the parity's field of the reference (row pitch and clamp width 2 wc,
        // see luma_list) and the vertical chroma vector offset of Table 8-10
        // (H264InterPrediction.cpp:2024-2043)
        const int fw = wc * g.step, fh = hc / g.step;
        int mvy0 = mb->mv[0][b][1], mvy1 = mb->mv[1][b][1];
        uint64_t o0 = coff, o1 = coff;
        if (g.field) {
            const int par0 = (ax.fpar >> q) & 1, par1 = (ax.fpar >> (4 + q)) & 1;
            o0 += par0 ? wc : 0;
            o1 += par1 ? wc : 0;
            mvy0 += par0 == g.bot ? 0 : (par0 ? -2 : 2);
            mvy1 += par1 == g.bot ? 0 : (par1 ? -2 : 2);
        }
        if (pf0) chroma2(slot_base(P, ax.refid[0][q]) + o0, fw, fh, xC, yC, mb->mv[0][b][0], mvy0, c00, c01);
        if (pf1) chroma2(slot_base(P, ax.refid[1][q]) + o1, fw, fh, xC, yC, mb->mv[1][b][0], mvy1, c10, c11);
        const DevWeight &w = G(P.wt)[ax.wp[q]];
        uint8_t *plane = c == 0 ? dcb : dcr;
        int v0, v1;
        if (!w.mode) {
            v0 = pf0 && pf1 ? (c00 + c10 + 1) >> 1 : (pf0 ? c00 : c10);
            v1 = pf0 && pf1 ? (c01 + c11 + 1) >> 1 : (pf0 ? c01 : c11);
        } else {
            const WpTerms t = wp_terms(w, 1 + c, pf0, pf1);
            v0 = wp_apply(t, (uint32_t)c00 | (uint32_t)c10 << 16);
            v1 = wp_apply(t, (uint32_t)c01 | (uint32_t)c11 << 16);
        }
        if (res) {
            v0 = clip1(v0 + ws->rc[c][(by * 2 + r) * 8 + bx * 2]);
            v1 = clip1(v1 + ws->rc[c][(by * 2 + r) * 8 + bx * 2 + 1]);
        }
        *(uint16_t *)&plane[(uint64_t)(g.cy + (by * 2 + r) * g.step) * wc + xC] = (uint16_t)(v0 | (v1 << 8));
    }
    (void)H;
}

// ------------------------------------------------------------ wavefront helpers
__device__ __forceinline__ uint32_t ld_acquire(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr uint32_t SPIN_LIMIT = 1u << 20;

// The persistent loops below keep all control flow around the ticket, the waits and
// the publication wave-uniform: every lane executes the same atomics, loads and stores
// (lane-identical addresses and values), and each value a branch depends on is made
// scalar with readfirstlane.  Lane-conditional regions here (e.g. "lane 0 publishes,
// then lane 0 takes the next ticket") let the compiler thread the two regions together
// across the loop latch; its CFG structurizer then leaves lanes 1..63 iterating on
// their own, which never terminates.
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Waits until flags[i] == epoch for each listed neighbour.  Bounded: after SPIN_LIMIT
// polls, or once another wave has raised the error word, the wave stops waiting (its
// output is then wrong, and h264g_sync reports the error).
// One of them (d < 0: none); the caller issues the acquire fence.
__device__ __forceinline__ void wait_dep(const uint32_t *flags, int dep, uint32_t epoch, uint32_t *err) {
    const int d = (int)uniform((uint32_t)dep);
    if (d < 0) return;
    for (uint32_t it = 0;; ++it) {
        if (uniform(ld_acquire(&flags[d])) == epoch) break;
        if (it >= SPIN_LIMIT) {
            atomicOr(err, DEVERR_TIMEOUT);
            break;
        }
        if ((it & 255u) == 255u && uniform(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0u) break;
        __builtin_amdgcn_s_sleep(1);
    }
}
__device__ void wait_deps(const uint32_t *flags, const int *deps, int nd, uint32_t epoch, uint32_t *err) {
    for (int k = 0; k < nd; k++) wait_dep(flags, deps[k], epoch, err);
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
}

// The barrier orders every lane's stores before the (wave-wide) release store.
__device__ void publish(uint32_t *flag, uint32_t epoch) {
    wave_sync();
    st_release(flag, epoch);
}

// Next ticket of a persistent launch (same value in every lane), or -1.  Only lane 0
// adds 1; the compiler's wave-level atomic combining issues one atomic for the wave.
__device__ int next_ticket(DevCtl *ctl, uint32_t n) {
    const uint32_t t = uniform(atomicAdd(&ctl->ticket, wave_lane() == 0 ? 1u : 0u));
    return t < n ? (int)t : -1;
}

// The launch's epoch: progress words equal to it are this launch's.  Read once at
// kernel start; it only changes when the launch's last wave retires, after every
// wave of the launch has read it.
__device__ __forceinline__ uint32_t launch_epoch(const uint32_t *epoch) {
    return uniform(__hip_atomic_load(epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// The last wave of a launch resets the control block for the next launch on the stream
// and advances the shared epoch, so the next wavefront launch (of either kind) uses a
// value no progress word holds yet.  Launch arguments therefore stay identical from one
// launch to the next, which lets a captured graph replay them.
__device__ void retire(DevCtl *ctl, uint32_t total_waves, uint32_t *epoch, uint32_t cur) {
    const uint32_t d = uniform(atomicAdd(&ctl->done, wave_lane() == 0 ? 1u : 0u));
    if (d == total_waves - 1) {
        __hip_atomic_store(&ctl->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(epoch, cur + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ __forceinline__ bool cls_inter(int c) { return c == H264G_MB_P || c == H264G_MB_P_SKIP || c == H264G_MB_B; }
// An MB that constrained intra prediction hides from intra MBs: the reference tests
// IS_INTER_Prediction_Mode(m_mb_pred_mode) (H264PictureBase.cpp:1129; H264CommonFunc.h:273),
// false for P_8x8 / B_8x8 (Pred_NA, H264MacroBlock.cpp:99-100, 142), which stay available.
__device__ __forceinline__ bool cip_hides(const h264g_mb &m) {
    return m.mb_type == H264G_MB_P_SKIP ||
           ((m.mb_type == H264G_MB_P || m.mb_type == H264G_MB_B) && m.part_shape != H264G_PART_8x8);
}
__device__ __forceinline__ bool cls_intra(int c) { return c == H264G_MB_I4x4 || c == H264G_MB_I8x8 || c == H264G_MB_I16x16; }

// ------------------------------------------------------------ intra (8.3)
// LDS tile of one wave: luma T[y+1][x+1] for x in [-1, 23], y in [-1, 15] (row -1 holds
// the top and top-right neighbours, column -1 the left ones), chroma C[c][y+1][x+1] for
// x, y in [-1, 7]; availability of the neighbouring MBs A, B, C, D.
struct WaveIntra {
    WaveRes res;
    uint8_t T[17][25];
    uint8_t C[2][9][9];
    int16_t pf[25];        // filtered 8x8 reference samples p'[-1,-1..7] then p'[0..15,-1]
    int16_t U[25];         // the same samples before filtering (-1: not available)
    int16_t X[2][16];      // neighbour lines of two 4x4 blocks (see pred4x4)
    uint8_t cbot[2][8];    // the MB's bottom chroma rows (for the intra strips)
    union {
        uint32_t p4[10][WAVE];  // Intra_4x4 wavefront step t, lane: see intra4_prep
        struct {
            int32_t ua[4][32];      // Intra_8x8 block, reference sample: see intra8_prep
            uint32_t tp[4][WAVE];   // Intra_8x8 block, pixel: taps on the filtered samples
        } p8;
    };
};

// Availability of the neighbouring MBs A, B, C, D for intra prediction as bits 0..3 of
// one word (a struct of four fields selected per lane would be indexed in scratch).
struct Avail {
    uint32_t m;
    __device__ bool A() const { return m & 1u; }
    __device__ bool B() const { return m & 2u; }
    __device__ bool C() const { return m & 4u; }
    __device__ bool D() const { return m & 8u; }
};

// Flat index into the tile WaveIntra::T (T[0][0] = p[-1, -1] of the MB, T[1 + y][1 + x] =
// sample (x, y)) of sample p[x, y] relative to the 4x4 / 8x8 block at (xO, yO) of the
// current MB, or -1 when it is not available for intra prediction (6.4.12.1,
// H264PictureBase.cpp:1099-1155; blk4: the 4x4 block, whose top-right samples inside the
// MB are not available for blocks 3 and 11, H264PictureBase.cpp:1131).
__device__ __forceinline__ int tile_a(const Avail &av, int xO, int yO, int x, int y, int blk4) {
    const int xx = xO + x, yy = yO + y;
    if (yy < 0) {
        if (xx < 0) return av.D() ? 0 : -1;
        if (xx < 16) return av.B() ? xx + 1 : -1;
        return av.C() ? xx + 1 : -1;
    }
    if (xx < 0) return av.A() ? (yy + 1) * 25 : -1;
    if (xx >= 16) return -1;
    if (blk4 >= 0 && x > 3 && (blk4 == 3 || blk4 == 11)) return -1;
    return (yy + 1) * 25 + xx + 1;
}

// 8.3.1.2 for pixel (x, y) of a 4x4 block (H264PictureBase.cpp:1062-1419), from the
// block's neighbour line X[0..14] = L3 L3 L2 L1 L0 TL T0 .. T7 T7 (L = p[-1, y],
// T = p[x, -1], TL = p[-1, -1]; T4..T7 already substituted by T3 when unavailable;
// -1 = not available).  Every directional mode reads a copy X[i], a 2-tap
// (X[i] + X[i+1] + 1) >> 1 or a 3-tap (X[c-1] + 2 X[c] + X[c+1] + 2) >> 2 of this line:
// the equations of 8.3.1.2.4 - 8.3.1.2.9 with p[-1, y] and p[x, -1] mapped to their X
// positions (the duplicated end samples give the corner cases, e.g. Diagonal_Down_Left
// at (3, 3) and Horizontal_Up at zHU = 5).  All three are (X[i0] + X[i1] + X[i2] + X[i3]
// + 2) >> 2: a copy repeats one index four times, a 2-tap equals (2 X[i] + 2 X[i+1] + 2)
// >> 2 and a 3-tap repeats its middle index.  i4_taps gives the four indices (4 bits each,
// i0 lowest) of sample (x, y) in mode m; DC (mode 2) is 0xFFFF.
constexpr uint16_t i4_copy(int i) { return (uint16_t)(i | i << 4 | i << 8 | i << 12); }
constexpr uint16_t i4_tap2(int i) { return (uint16_t)(i | i << 4 | (i + 1) << 8 | (i + 1) << 12); }
constexpr uint16_t i4_tap3(int c) { return (uint16_t)((c - 1) | c << 4 | c << 8 | (c + 1) << 12); }
constexpr uint16_t i4_taps(int mode, int x, int y) {
    switch (mode) {
    case 0: return i4_copy(6 + x);
    case 1: return i4_copy(4 - y);
    case 2: return 0xFFFF;
    case 3: return i4_tap3(7 + x + y);
    case 4: return i4_tap3(5 + x - y);
    case 5: {
        const int z = 2 * x - y;
        if (z >= 0) return (z & 1) ? i4_tap3(5 + x - (y >> 1)) : i4_tap2(5 + x - (y >> 1));
        return z == -1 ? i4_tap3(5) : i4_tap3(6 - y);
    }
    case 6: {
        const int z = 2 * y - x;
        if (z >= 0) return (z & 1) ? i4_tap3(5 - y + (x >> 1)) : i4_tap2(4 - y + (x >> 1));
        return z == -1 ? i4_tap3(5) : i4_tap3(4 + x);
    }
    case 7: return (y & 1) ? i4_tap3(7 + x + (y >> 1)) : i4_tap2(6 + x + (y >> 1));
    default: {
        const int z = x + 2 * y;
        if (z < 5) return (z & 1) ? i4_tap3(3 - y - (x >> 1)) : i4_tap2(3 - y - (x >> 1));
        return z == 5 ? i4_tap3(1) : i4_copy(1);
    }
    }
}
struct I4Taps {
    uint16_t v[9][16];
};
constexpr I4Taps make_i4taps() {
    I4Taps t{};
    for (int m = 0; m < 9; m++)
        for (int k = 0; k < 16; k++) t.v[m][k] = i4_taps(m, k & 3, k >> 2);
    return t;
}
__constant__ I4Taps c_i4taps = make_i4taps();

// What an Intra_4x4 MB's wavefront steps need that does not depend on samples, gathered
// into wi->p4 before the MB waits for its neighbours: per step t and lane the flat tile
// index of neighbour-line sample j (lanes j < 15 of each 16-lane group; -1: not
// available, T4..T7 already substituted by T3) in the low half and the taps of pixel j of
// the lane's block (c_i4taps) in the high half.  The LDS writes are ordered before the
// MB's first wave barrier in intra_mb.
__device__ __forceinline__ void intra4_prep(const h264g_mb *mb, const Avail &av, WaveIntra *wi) {
    const int lane = wave_lane(), g = lane >> 4, j = lane & 15;
#pragma unroll 1
    for (int t = 0; t < 10; t++) {
        const int by = (t < 2 ? 0 : (t - 2) >> 1) + g, bx = t - 2 * by;
        const bool has = lane < 32 && bx >= 0 && by < 4;
        const int blk = 8 * (by >> 1) + 4 * (bx >> 1) + 2 * (by & 1) + (bx & 1), xO = bx * 4, yO = by * 4;
        int a = -1;
        uint32_t tp = 0;
        if (has) {
            if (j < 5) a = tile_a(av, xO, yO, -1, j == 0 ? 3 : 4 - j, blk);
            else if (j == 5) a = tile_a(av, xO, yO, -1, -1, blk);
            else if (j < 15) {
                const int tx = j == 14 ? 7 : j - 6;
                a = tile_a(av, xO, yO, tx, -1, blk);
                if (tx >= 4 && a < 0) {  // p[4..7, -1] <- p[3, -1] (H264PictureBase.cpp:1161-1167)
                    const int a3 = tile_a(av, xO, yO, 3, -1, blk);
                    if (a3 >= 0) a = a3;
                }
            }
            tp = c_i4taps.v[mb->intra_pred_modes[blk]][j];
        }
        wi->p4[t][lane] = ((uint32_t)a & 0xFFFFu) | tp << 16;
    }
}

// Intra_8x8 sample (x, y) of mode m (8.3.2.2.2 - 8.3.2.2.10, H264PictureBase.cpp:1601-1843)
// as (pf[i0] + pf[i1] + pf[i2] + pf[i3] + 2) >> 2 on the filtered reference samples
// pf[0] = p'[-1, -1], pf[1 + y] = p'[-1, y], pf[9 + x] = p'[x, -1], like i4_taps (copy =
// one index four times, 2-tap = both twice, 3-tap = middle twice; a weight 3 repeats an
// index three times); 5 bits each, i0 lowest; DC (mode 2) is 0xFFFFFFFF.
constexpr int i8_ft(int i) { return i < 0 ? 0 : 9 + i; }
constexpr int i8_fl(int i) { return i < 0 ? 0 : 1 + i; }
constexpr uint32_t i8_pack(int a, int b, int c, int d) { return (uint32_t)(a | b << 5 | c << 10 | d << 15); }
constexpr uint32_t i8_tap2(int a, int b) { return i8_pack(a, a, b, b); }
constexpr uint32_t i8_tap3(int a, int c, int b) { return i8_pack(a, c, c, b); }
constexpr uint32_t i8_taps(int mode, int x, int y) {
    switch (mode) {
    case 0: return i8_pack(i8_ft(x), i8_ft(x), i8_ft(x), i8_ft(x));
    case 1: return i8_pack(i8_fl(y), i8_fl(y), i8_fl(y), i8_fl(y));
    case 2: return 0xFFFFFFFFu;
    case 3:
        if (x == 7 && y == 7) return i8_pack(i8_ft(14), i8_ft(15), i8_ft(15), i8_ft(15));
        return i8_tap3(i8_ft(x + y), i8_ft(x + y + 1), i8_ft(x + y + 2));
    case 4:
        if (x > y) return i8_tap3(i8_ft(x - y - 2), i8_ft(x - y - 1), i8_ft(x - y));
        if (x < y) return i8_tap3(i8_fl(y - x - 2), i8_fl(y - x - 1), i8_fl(y - x));
        return i8_tap3(i8_ft(0), 0, i8_fl(0));
    case 5: {
        const int z = 2 * x - y;
        if (z >= 0 && !(z & 1)) return i8_tap2(i8_ft(x - (y >> 1) - 1), i8_ft(x - (y >> 1)));
        if (z >= 0) return i8_tap3(i8_ft(x - (y >> 1) - 2), i8_ft(x - (y >> 1) - 1), i8_ft(x - (y >> 1)));
        if (z == -1) return i8_tap3(i8_fl(0), 0, i8_ft(0));
        return i8_tap3(i8_fl(y - 2 * x - 1), i8_fl(y - 2 * x - 2), i8_fl(y - 2 * x - 3));
    }
    case 6: {
        const int z = 2 * y - x;
        if (z >= 0 && !(z & 1)) return i8_tap2(i8_fl(y - (x >> 1) - 1), i8_fl(y - (x >> 1)));
        if (z >= 0) return i8_tap3(i8_fl(y - (x >> 1) - 2), i8_fl(y - (x >> 1) - 1), i8_fl(y - (x >> 1)));
        if (z == -1) return i8_tap3(i8_fl(0), 0, i8_ft(0));
        return i8_tap3(i8_ft(x - 2 * y - 1), i8_ft(x - 2 * y - 2), i8_ft(x - 2 * y - 3));
    }
    case 7:
        if (!(y & 1)) return i8_tap2(i8_ft(x + (y >> 1)), i8_ft(x + (y >> 1) + 1));
        return i8_tap3(i8_ft(x + (y >> 1)), i8_ft(x + (y >> 1) + 1), i8_ft(x + (y >> 1) + 2));
    default: {
        const int z = x + 2 * y;
        if (z < 13 && !(z & 1)) return i8_tap2(i8_fl(y + (x >> 1)), i8_fl(y + (x >> 1) + 1));
        if (z < 13) return i8_tap3(i8_fl(y + (x >> 1)), i8_fl(y + (x >> 1) + 1), i8_fl(y + (x >> 1) + 2));
        if (z == 13) return i8_pack(i8_fl(6), i8_fl(7), i8_fl(7), i8_fl(7));
        return i8_pack(i8_fl(7), i8_fl(7), i8_fl(7), i8_fl(7));
    }
    }
}
struct I8Taps {
    uint32_t v[9][64];
};
constexpr I8Taps make_i8taps() {
    I8Taps t{};
    for (int m = 0; m < 9; m++)
        for (int k = 0; k < 64; k++) t.v[m][k] = i8_taps(m, k & 7, k >> 3);
    return t;
}
__constant__ I8Taps c_i8taps = make_i8taps();

// What an Intra_8x8 MB's four blocks need that does not depend on samples, gathered into
// wi->p8 before the MB waits for its neighbours: per block the flat tile index of each of
// the 25 reference samples before filtering (U[0] = p[-1, -1], U[1 + y] = p[-1, y],
// U[9 + x] = p[x, -1] with p[8..15, -1] <- p[7, -1] when not available; -1: not
// available) and each pixel's taps (c_i8taps).
__device__ __forceinline__ void intra8_prep(const h264g_mb *mb, const Avail &av, WaveIntra *wi) {
    const int lane = wave_lane();
#pragma unroll 1
    for (int blk = 0; blk < 4; blk++) {
        const int xO = (blk & 1) * 8, yO = (blk >> 1) * 8;
        if (lane < 25) {
            int a;
            if (lane == 0) a = tile_a(av, xO, yO, -1, -1, -1);
            else if (lane < 9) a = tile_a(av, xO, yO, -1, lane - 1, -1);
            else {
                const int x = lane - 9;
                a = tile_a(av, xO, yO, x, -1, -1);
                if (x >= 8 && a < 0) {
                    const int a7 = tile_a(av, xO, yO, 7, -1, -1);
                    if (a7 >= 0) a = a7;
                }
            }
            wi->p8.ua[blk][lane] = a;
        }
        wi->p8.tp[blk][lane] = c_i8taps.v[mb->intra_pred_modes[blk]][lane];
    }
}

// Residuals of intra MBs.  They need no neighbour samples, so the inter stage, which is
// parallel over MBs, builds them (intra_res_build) and leaves them in G(P.ires); the intra
// schedules load an MB's residual (intra_res_fetch, lanes 0-47 one 16-byte piece each)
// before they wait for its neighbours and put it in the tile after the wait
// (intra_res_place).
static_assert(offsetof(WaveRes, rc) == offsetof(WaveRes, rl) + 512 && offsetof(WaveRes, rl) % 16 == 0,
              "luma and chroma residual planes are one 768-byte run");
template <bool MBAFF, bool SCALED>
__device__ __forceinline__ void intra_res_build(const DevPic &P, int addr, WaveRes *ws) {
    const DevAux &ax = G(P.aux)[addr];
    if (!(SCALED && P.ls ? build_residual<MBAFF, SCALED>(&G(P.mbs)[addr], ax, G(P.levels), ws, G(P.ls))
                         : build_residual<MBAFF, false>(&G(P.mbs)[addr], ax, G(P.levels), ws, nullptr)))
        return;
    const int lane = wave_lane();
    if (lane < 48) ((uint4 *)(G(P.ires) + (size_t)(ax.ires & AUX_IRES_MASK) * 384))[lane] = ((const uint4 *)ws->rl)[lane];
}
struct IntraRes {
    uint4 v;
    bool res;
};
__device__ __forceinline__ IntraRes intra_res_fetch(const DevPic &P, int addr) {
    const DevAux &ax = G(P.aux)[addr];
    IntraRes r;
    r.res = (ax.kind & AUX_RESID) != 0;
    r.v = make_uint4(0, 0, 0, 0);
    if (r.res && wave_lane() < 48) r.v = ((const uint4 *)(G(P.ires) + (size_t)(ax.ires & AUX_IRES_MASK) * 384))[wave_lane()];
    return r;
}
__device__ __forceinline__ void intra_res_place(const IntraRes &r, WaveIntra *wi) {
    if (r.res && wave_lane() < 48) ((uint4 *)wi->res.rl)[wave_lane()] = r.v;
}

// Prediction and reconstruction of an intra MB whose residual intra_res_place left in
// wi->res (res: the MB has levels); the neighbours it reads are final.  The LDS writes of
// the residual are ordered before its first wave barrier.
// above: neighbours (Avail bits: 2 = B, 4 = C, 8 = D) whose samples of row -1 the caller
// already put into the tile, from the intra strips.
// An Intra_4x4 MB has run intra4_prep, an Intra_8x8 MB intra8_prep.
template <bool MBAFF>
__device__ __forceinline__ void intra_mb(const DevPic &P, int addr, WaveIntra *wi, const Avail &av, bool res,
                                         uint32_t above = 0) {
    const int lane = wave_lane();
    const h264g_mb *mb = &G(P.mbs)[addr];
    const int W = P.W, H = P.H, wl = W * 16, hl = H * 16, wc = W * 8, hc = H * 8;
    // In an MBAFF frame every neighbouring sample that 6.4.12.2 (Table 6-4) gives an MB lies
    // in the MB's own view of the picture: frame rows for frame MBs, the rows of its parity
    // for field MBs (rows step apart from its first row), at the same columns; Table 6-4
    // only decides availability (the caller's av).
    const MbGeo g = mb_geo<MBAFF>(P, addr);
    const int mbx = g.mbx, st = g.step;
    uint8_t *dst = slot_base(P, P.dst);
    uint8_t *dy = dst, *dcb = dst + (uint64_t)wl * hl, *dcr = dcb + (uint64_t)wc * hc;
    // ---- neighbours into the tile
    if (lane < 25) {  // row -1, x = lane - 1
        int x = lane - 1;
        bool ok = x < 0 ? av.D() : (x < 16 ? av.B() : av.C());
        ok = ok && !(above & (x < 0 ? 8u : (x < 16 ? 2u : 4u)));
        if (ok) wi->T[0][lane] = dy[(uint64_t)(g.ly - st) * wl + mbx * 16 + x];
    } else if (lane < 41) {  // column -1
        int y = lane - 25;
        if (av.A()) wi->T[y + 1][0] = dy[(uint64_t)(g.ly + y * st) * wl + mbx * 16 - 1];
    } else if (lane < 59) {  // chroma row -1 (x = -1..7) for both planes
        int k = lane - 41, c = k / 9, x = k % 9 - 1;
        bool ok = (x < 0 ? av.D() : av.B()) && !(above & (x < 0 ? 8u : 2u));
        const uint8_t *pl = c == 0 ? dcb : dcr;
        if (ok) wi->C[c][0][x + 1] = pl[(uint64_t)(g.cy - st) * wc + mbx * 8 + x];
    }
    if (lane < 16) {  // chroma column -1
        int c = lane >> 3, y = lane & 7;
        const uint8_t *pl = c == 0 ? dcb : dcr;
        if (av.A()) wi->C[c][y + 1][0] = pl[(uint64_t)(g.cy + y * st) * wc + mbx * 8 - 1];
    }
    wave_sync();
    // ---- luma
    if (mb->mb_type == H264G_MB_I4x4) {
        // 4x4 blocks in wavefront order t = bx + 2 by, two blocks per step (lanes 0-15 and
        // 16-31): every neighbour a block reads has a smaller t, and a top-right block that
        // the reference's block order has not decoded yet (blocks 3, 11) counts as not
        // available by rule (tile_a), so the result equals the block-index order.  Each
        // step gathers the two neighbour lines X (pred4x4) from the tile at the indices
        // intra4_prep found, then evaluates every pixel with its four taps.
        const int g = lane >> 4, j = lane & 15, x = j & 3, y = j >> 2;
        const uint8_t *Tf = &wi->T[0][0];
        const int16_t *X = wi->X[g];
#pragma unroll 1
        for (int t = 0; t < 10; t++) {
            const int by = (t < 2 ? 0 : (t - 2) >> 1) + g, bx = t - 2 * by;
            const bool has = lane < 32 && bx >= 0 && by < 4;
            const int xO = bx * 4, yO = by * 4;
            const uint32_t e = wi->p4[t][lane];
            const int xa = (int)(int16_t)(e & 0xFFFFu);
            const uint32_t tp = e >> 16;
            if (has && j < 15) wi->X[g][j] = (int16_t)(xa >= 0 ? (int)Tf[xa] : -1);
            wave_sync();
            if (has) {
                int v = (X[tp & 15u] + X[(tp >> 4) & 15u] + X[(tp >> 8) & 15u] + X[tp >> 12] + 2) >> 2;
                if (tp == 0xFFFFu) {  // DC: the top / left samples are available all or none
                    const bool ta = X[6] >= 0, la = X[4] >= 0;
                    const int st = X[6] + X[7] + X[8] + X[9], sl = X[1] + X[2] + X[3] + X[4];
                    v = ta && la ? (st + sl + 4) >> 3 : la ? (sl + 2) >> 2 : ta ? (st + 2) >> 2 : 128;
                }
                if (res) v = clip1(v + wi->res.rl[(yO + y) * 16 + xO + x]);
                wi->T[yO + y + 1][xO + x + 1] = (uint8_t)v;
            }
            wave_sync();
        }
    } else if (mb->mb_type == H264G_MB_I8x8) {
        // each block: gather its 25 reference samples at the indices intra8_prep found,
        // filter them, then one four-tap sum per pixel
        const uint8_t *Tf = &wi->T[0][0];
#pragma unroll 1
        for (int blk = 0; blk < 4; blk++) {
            const int xO = (blk & 1) * 8, yO = (blk >> 1) * 8;
            if (lane < 25) {
                const int a = wi->p8.ua[blk][lane];
                wi->U[lane] = (int16_t)(a >= 0 ? (int)Tf[a] : -1);
            }
            wave_sync();
            // 8.3.2.2.1 reference sample filtering (H264PictureBase.cpp:1535-1599); the left
            // column and the 16 top samples are available all or none
            if (lane < 25) {
                const int16_t *U = wi->U;
                const int pm = U[0];
                int v = -1;
                if (lane == 0) {  // p'[-1, -1]
                    const int pt = U[9], pl = U[1];
                    if (pm >= 0) {
                        if (pt < 0 || pl < 0) v = pt >= 0 ? (3 * pm + pt + 2) >> 2 : (pl >= 0 ? (3 * pm + pl + 2) >> 2 : pm);
                        else v = (pt + 2 * pm + pl + 2) >> 2;
                    }
                } else if (lane < 9) {  // p'[-1, y]
                    const int y = lane - 1;
                    if (U[1] >= 0) {
                        if (y == 0) v = pm >= 0 ? (pm + 2 * U[1] + U[2] + 2) >> 2 : (3 * U[1] + U[2] + 2) >> 2;
                        else if (y == 7) v = (U[7] + 3 * U[8] + 2) >> 2;
                        else v = (U[y] + 2 * U[y + 1] + U[y + 2] + 2) >> 2;
                    }
                } else {  // p'[x, -1]
                    const int x = lane - 9;
                    if (U[9] >= 0 && U[17] >= 0) {
                        if (x == 0) v = pm >= 0 ? (pm + 2 * U[9] + U[10] + 2) >> 2 : (3 * U[9] + U[10] + 2) >> 2;
                        else if (x == 15) v = (U[23] + 3 * U[24] + 2) >> 2;
                        else v = (U[8 + x] + 2 * U[9 + x] + U[10 + x] + 2) >> 2;
                    }
                }
                wi->pf[lane] = (int16_t)v;
            }
            wave_sync();
            const int x = lane & 7, y = lane >> 3;
            const uint32_t tp = wi->p8.tp[blk][lane];
            const int16_t *pf = wi->pf;
            int v = (pf[tp & 31u] + pf[(tp >> 5) & 31u] + pf[(tp >> 10) & 31u] + pf[(tp >> 15) & 31u] + 2) >> 2;
            if (tp == 0xFFFFFFFFu) {  // DC (8.3.2.2.4): the filtered samples are available all or none
                const bool ta = pf[9] >= 0, la = pf[1] >= 0;
                int st = 0, sl = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) { st += pf[9 + k]; sl += pf[1 + k]; }
                v = ta && la ? (st + sl + 8) >> 4 : la ? (sl + 4) >> 3 : ta ? (st + 4) >> 3 : 128;
            }
            if (res) v = clip1(v + wi->res.rl[(yO + y) * 16 + xO + x]);
            wave_sync();
            wi->T[yO + y + 1][xO + x + 1] = (uint8_t)v;
            wave_sync();
        }
    } else {  // Intra_16x16, 8.3.3 (H264PictureBase.cpp:1847-2056)
        int mode = mb->intra16x16_pred_mode;
        int dc = 128, a = 0, bq = 0, cq = 0;
        if (mode == 2) {
            int st = 0, sl = 0;
            for (int k = 0; k < 16; k++) { st += wi->T[0][k + 1]; sl += wi->T[k + 1][0]; }
            if (av.A() && av.B()) dc = (st + sl + 16) >> 5;
            else if (av.A()) dc = (sl + 8) >> 4;
            else if (av.B()) dc = (st + 8) >> 4;
        } else if (mode == 3) {
            int Hs = 0, Vs = 0;
            for (int k = 0; k < 8; k++) {
                Hs += (k + 1) * (wi->T[0][8 + k + 1] - wi->T[0][6 - k + 1]);
                Vs += (k + 1) * (wi->T[8 + k + 1][0] - wi->T[6 - k + 1][0]);
            }
            a = 16 * (wi->T[16][0] + wi->T[0][16]);
            bq = (5 * Hs + 32) >> 6;
            cq = (5 * Vs + 32) >> 6;
        }
        for (int k = 0; k < 4; k++) {
            int i = lane + k * WAVE, x = i & 15, y = i >> 4;
            int v = mode == 0 ? wi->T[0][x + 1] : mode == 1 ? wi->T[y + 1][0] : mode == 2 ? dc : clip1((a + bq * (x - 7) + cq * (y - 7) + 16) >> 5);
            if (res) v = clip1(v + wi->res.rl[y * 16 + x]);
            wi->T[y + 1][x + 1] = (uint8_t)v;  // only this lane reads/writes pixel i
        }
        wave_sync();
    }
    // ---- chroma, 8.3.4 (H264PictureBase.cpp:2076-2381); DC tests "> 0" (quirk Q1)
    {
        int mode = mb->intra_chroma_pred_mode;
        for (int k = 0; k < 2; k++) {
            int i = lane + k * WAVE, c = i >> 6, x = i & 7, y = (i >> 3) & 7;
            const uint8_t (*Cp)[9] = wi->C[c];
            int v;
            if (mode == 0) {
                int xO = x & 4, yO = y & 4;
                int t = av.B() && Cp[0][xO + 1] > 0 && Cp[0][xO + 2] > 0 && Cp[0][xO + 3] > 0 && Cp[0][xO + 4] > 0;
                int l = av.A() && Cp[yO + 1][0] > 0 && Cp[yO + 2][0] > 0 && Cp[yO + 3][0] > 0 && Cp[yO + 4][0] > 0;
                int st = Cp[0][xO + 1] + Cp[0][xO + 2] + Cp[0][xO + 3] + Cp[0][xO + 4];
                int sl = Cp[yO + 1][0] + Cp[yO + 2][0] + Cp[yO + 3][0] + Cp[yO + 4][0];
                if ((xO == 0 && yO == 0) || (xO > 0 && yO > 0)) v = (t && l) ? (st + sl + 4) >> 3 : (l ? (sl + 2) >> 2 : (t ? (st + 2) >> 2 : 128));
                else if (xO > 0) v = t ? (st + 2) >> 2 : (l ? (sl + 2) >> 2 : 128);
                else v = l ? (sl + 2) >> 2 : (t ? (st + 2) >> 2 : 128);
            } else if (mode == 1) {
                v = Cp[y + 1][0];
            } else if (mode == 2) {
                v = Cp[0][x + 1];
            } else {
                int Hs = 0, Vs = 0;
                for (int j = 0; j < 4; j++) {
                    Hs += (j + 1) * (Cp[0][4 + j + 1] - Cp[0][2 - j + 1]);
                    Vs += (j + 1) * (Cp[4 + j + 1][0] - Cp[2 - j + 1][0]);
                }
                int aa = 16 * (Cp[8][0] + Cp[0][8]), bb = (34 * Hs + 32) >> 6, cc = (34 * Vs + 32) >> 6;
                v = clip1((aa + bb * (x - 3) + cc * (y - 3) + 16) >> 5);
            }
            if (res) v = clip1(v + wi->res.rc[c][y * 8 + x]);
            if (y == 7) wi->cbot[c][x] = (uint8_t)v;
            uint8_t *pl = c == 0 ? dcb : dcr;
            pl[(uint64_t)(g.cy + y * st) * wc + mbx * 8 + x] = (uint8_t)v;
        }
    }
    // ---- luma tile -> picture
    for (int k = 0; k < 4; k++) {
        int i = lane + k * WAVE, x = i & 15, y = i >> 4;
        dy[(uint64_t)(g.ly + y * st) * wl + mbx * 16 + x] = wi->T[y + 1][x + 1];
    }
    (void)H;
}

// ------------------------------------------------------------ deblocking (8.7)
// What the bS derivation reads of an MB.  Scalar fields only, so that selecting between
// views and per-lane field choices stay in registers.
struct MbView {
    int cls, kind, nz4, nz8;
    uint32_t ref0, ref1;  // reference DPB slot per quadrant q (int8 in byte q), lists 0 / 1
    int pf;
};
__device__ __forceinline__ int view_ref(const MbView &v, int l, int q) { return (int8_t)((l ? v.ref1 : v.ref0) >> (8 * q)); }
__device__ __forceinline__ int dbq_qp(uint32_t q, int plane) { return (q >> (8 * plane)) & 255; }
__device__ __forceinline__ bool dbq_t8(uint32_t q) { return (q >> 24) & 1u; }
// the MB's left / top MB edge lies between slices under disable_deblocking_filter_idc 0:
// its p side takes the MB's own qP (see k_bs)
constexpr uint32_t DBQ_OWN_LEFT = 1u << 25, DBQ_OWN_TOP = 1u << 26;

__device__ __forceinline__ void load_view(const DevPic &P, int addr, MbView &v) {
    const DevAux ax = G(P.aux)[addr];
    v.cls = G(P.mbs)[addr].mb_type;
    v.kind = ax.kind;
    v.nz4 = ax.nz4;
    v.nz8 = ax.nz8;
    v.pf = ax.pf;
    v.ref0 = (uint32_t)(uint8_t)ax.refid[0][0] | ((uint32_t)(uint8_t)ax.refid[0][1] << 8) |
             ((uint32_t)(uint8_t)ax.refid[0][2] << 16) | ((uint32_t)(uint8_t)ax.refid[0][3] << 24);
    v.ref1 = (uint32_t)(uint8_t)ax.refid[1][0] | ((uint32_t)(uint8_t)ax.refid[1][1] << 8) |
             ((uint32_t)(uint8_t)ax.refid[1][2] << 16) | ((uint32_t)(uint8_t)ax.refid[1][3] << 24);
}
// Deblocking word of an MB (G(P.dbq), written by k_bs): qP of luma, Cb edges, Cr edges in
// bytes 0..2, transform_size_8x8_flag in bit 24, DBQ_OWN_LEFT / DBQ_OWN_TOP in bits 25, 26.

__device__ __forceinline__ bool mv_far(int ax, int ay, int bx, int by) { return iabs(ax - bx) >= 4 || iabs(ay - by) >= 4; }

// 8.7.2.1 (Deblock.cpp:994-1310) with quirk Q3; I_PCM counts as neither intra nor inter.
__device__ __forceinline__ int bs_luma(const DevPic &P, int ap, const MbView &Pv, int aq, const MbView &Qv, int mb_edge, int xp, int yp, int xq, int yq) {
    bool pin = Pv.kind & AUX_INTRA, qin = Qv.kind & AUX_INTRA;
    if (mb_edge && (pin || qin)) return 4;
    if (pin || qin) return 3;
    int bp4 = 8 * (yp >> 3) + 4 * (xp >> 3) + 2 * ((yp & 7) >> 2) + ((xp & 7) >> 2);
    int bq4 = 8 * (yq >> 3) + 4 * (xq >> 3) + 2 * ((yq & 7) >> 2) + ((xq & 7) >> 2);
    int bp8 = 2 * (yp >> 3) + (xp >> 3), bq8 = 2 * (yq >> 3) + (xq >> 3);
    bool pt8 = Pv.kind & AUX_T8, qt8 = Qv.kind & AUX_T8;
    if ((pt8 && ((Pv.nz8 >> bp8) & 1)) || (!pt8 && ((Pv.nz4 >> bp4) & 1)) ||
        (qt8 && ((Qv.nz8 >> bq8) & 1)) || (!qt8 && ((Qv.nz4 >> bq4) & 1)))
        return 2;
    int qp = 2 * (yp >> 3) + (xp >> 3), qq = 2 * (yq >> 3) + (xq >> 3);
    int rp = (yp >> 2) * 4 + (xp >> 2), rq = (yq >> 2) * 4 + (xq >> 2);
    int p0 = view_ref(Pv, 0, qp), p1 = view_ref(Pv, 1, qp), q0 = view_ref(Qv, 0, qq), q1 = view_ref(Qv, 1, qq);
    int fp0 = (Pv.pf >> qp) & 1, fp1 = (Pv.pf >> (4 + qp)) & 1, fq0 = (Qv.pf >> qq) & 1, fq1 = (Qv.pf >> (4 + qq)) & 1;
    if (!(((p0 == q0 && p1 == q1) || (p0 == q1 && p1 == q0)) && fp0 + fp1 == fq0 + fq1)) return 1;
    const h264g_mb *mp = &G(P.mbs)[ap], *mq = &G(P.mbs)[aq];
    int mp0x = mp->mv[0][rp][0], mp0y = mp->mv[0][rp][1], mp1x = mp->mv[1][rp][0], mp1y = mp->mv[1][rp][1];
    int mq0x = mq->mv[0][rq][0], mq0y = mq->mv[0][rq][1], mq1x = mq->mv[1][rq][0], mq1y = mq->mv[1][rq][1];
    if (fp0 && !fp1 && fq0 && !fq1 && mv_far(mp0x, mp0y, mq0x, mq0y)) return 1;
    if (fp0 && !fp1 && !fq0 && fq1 && mv_far(mp0x, mp0y, mq1x, mq1y)) return 1;
    if (!fp0 && fp1 && fq0 && !fq1 && mv_far(mp1x, mp1y, mq0x, mq0y)) return 1;
    if (!fp0 && fp1 && !fq0 && fq1 && mv_far(mp1x, mp1y, mq1x, mq1y)) return 1;
    if (fp0 && fp1 && p0 != p1 && fq0 && fq1 && ((q0 == p0 && q1 == p1) || (q0 == p1 && q1 == p0))) {
        if (q0 == p0 && (mv_far(mp0x, mp0y, mq0x, mq0y) || mv_far(mp1x, mp1y, mq1x, mq1y))) return 1;
        if (q0 == p1 && (mv_far(mp1x, mp1y, mq0x, mq0y) || mv_far(mp0x, mp0y, mq1x, mq1y))) return 1;
    }
    if (fp0 && fp1 && p0 == p1 && fq0 && fq1 && q0 == q1 && q0 == p0) {
        if (mv_far(mp0x, mp0y, mq0x, mq0y) || (mv_far(mp1x, mp1y, mq1x, mq1y) && mv_far(mp0x, mp0y, mq1x, mq1y)) ||
            mv_far(mp1x, mp1y, mq0x, mq0y))
            return 1;
    }
    return 0;
}

// Tables 8-16 (alpha', beta') and 8-17 (tC0), copied into LDS by every deblocking
// workgroup: lookups with lane-varying indices then cost an LDS read, not a vector
// memory access.
struct DbkTables {
    uint8_t alpha[52], beta[52], tc0[3][52];
};
__device__ void dbk_tables_init(DbkTables *t) {
    for (int i = threadIdx.x; i < 52; i += blockDim.x) {
        t->alpha[i] = c_alpha[i];
        t->beta[i] = c_beta[i];
        t->tc0[0][i] = c_tc0[0][i];
        t->tc0[1][i] = c_tc0[1][i];
        t->tc0[2][i] = c_tc0[2][i];
    }
}

// LDS working set of one deblocking wave, in MB-relative coordinates: luma rows [-4, 15]
// x columns [-16, 15] (the MB and the whole MB to its left), chroma rows [-4, 7] x
// columns [-8, 7].  Rows -4..-1 hold the bottom rows of the MB above.
constexpr int DT_LX = 16, DT_CX = 8;  // tile column of MB column 0 (luma, chroma)
struct alignas(16) WaveDbk {
    uint8_t L[20][32];
    uint8_t C[2][12][16];
};

// Plane pointers of the picture being deblocked.
struct DbkPlanes {
    uint8_t *y, *cb, *cr;
    int wl, wc;
};
__device__ __forceinline__ DbkPlanes dbk_planes(const DevPic &P) {
    DbkPlanes d;
    d.wl = P.W * 16;
    d.wc = P.W * 8;
    d.y = slot_base(P, P.dst);
    d.cb = d.y + (uint64_t)d.wl * (P.H * 16);
    d.cr = d.cb + (uint64_t)d.wc * (P.H * 8);
    return d;
}

// Tile loads.  Region bits: 1 = the MB itself (rows / cols >= 0), 2 = the 4 rows above
// it (cols >= 0), 4 = the 4 columns left of it (rows >= 0).  The corner is never used.
__device__ void dbk_load(const DbkPlanes &d, int mbx, int mby, int region, WaveDbk *wd) {
    const int lane = wave_lane();
    for (int i = lane; i < 400; i += WAVE) {
        const int y = i / 20 - 4, x = i % 20 - 4;
        const int r = (y >= 0 && x >= 0) ? 1 : (y < 0 && x >= 0 ? 2 : (x < 0 && y >= 0 ? 4 : 0));
        if (!(r & region)) continue;
        wd->L[y + 4][x + DT_LX] = d.y[(uint64_t)(mby * 16 + y) * d.wl + mbx * 16 + x];
    }
    for (int i = lane; i < 288; i += WAVE) {
        const int c = i / 144, k = i % 144, y = k / 12 - 4, x = k % 12 - 4;
        const int r = (y >= 0 && x >= 0) ? 1 : (y < 0 && x >= 0 ? 2 : (x < 0 && y >= 0 ? 4 : 0));
        if (!(r & region)) continue;
        const uint8_t *pl = c == 0 ? d.cb : d.cr;
        wd->C[c][y + 4][x + DT_CX] = pl[(uint64_t)(mby * 8 + y) * d.wc + mbx * 8 + x];
    }
}

__device__ __forceinline__ int absdiff(int a, int b) { return (int)__builtin_amdgcn_sad_u16((uint32_t)a, (uint32_t)b, 0u); }
// True when the condition holds on any active lane (the compare's lane mask, no VALU).
__device__ __forceinline__ bool any_lane(bool v) { return __builtin_amdgcn_ballot_w64(v) != 0; }

// Packed 16-bit pairs: the p-side value of a line in the low half, the q-side value in
// the high half.  The p and q sides of 8.7.2.3 / 8.7.2.4 are mirror images, so one packed
// integer operation (v_pk_*_i16) computes both.
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ s16x2 as_pair(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ uint32_t as_bits(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s16x2 pair(int lo, int hi) { return (s16x2){(short)lo, (short)hi}; }
__device__ __forceinline__ s16x2 pair_clip(s16x2 v, s16x2 lo, s16x2 hi) {
    return __builtin_elementwise_min(__builtin_elementwise_max(v, lo), hi);
}
__device__ __forceinline__ s16x2 pair_swap(s16x2 v) { return __builtin_shufflevector(v, v, 1, 0); }
// n where the half's condition holds (hp: p half, hq: q half), else o.
__device__ __forceinline__ s16x2 pair_sel(bool hp, bool hq, s16x2 n, s16x2 o) {
    const uint32_t m = (hp ? 0x0000ffffu : 0u) | (hq ? 0xffff0000u : 0u);
    return as_pair((as_bits(n) & m) | (as_bits(o) & ~m));
}

// Edge J of a line held as 20 bytes in w[0..4] (little-endian; the edge lies between
// words J and J + 1); chroma lines hold 12 bytes and use J = 0, 1 (luma edges 0 and 8
// for the bS).  8.7.2.2 - 8.7.2.4 (Deblock.cpp:1314-1522): alpha = alpha'(indexA), beta =
// beta'(indexB), tc0 = tC0'(bS, indexA) (used for bS 1..3), with indexA / indexB =
// Clip3(0, 51, qPav + filterOffsetA / B).  Every lane takes the same path: the wave skips
// a filter that no lane applies (wave-uniform branches); inside, every candidate value is
// computed on all lanes and chosen per lane with selects, so no per-line condition turns
// into a divergent branch and its exec-mask bookkeeping.  The samples travel as (p, q)
// pairs: P0 = (p0, q0) .. P3 = (p3, q3).
template <int J>
__device__ __forceinline__ void dbk_line_edge(uint32_t (&w)[5], bool act, bool luma, int bS, int alpha, int beta, int tc0) {
    const bool on = act && bS != 0;
    if (!any_lane(on)) return;
    const uint32_t a = w[J], b = w[J + 1];  // a: p3 p2 p1 p0 (bytes 0..3), b: q0 q1 q2 q3
    const s16x2 P0 = as_pair(__builtin_amdgcn_perm(b, a, 0x0c040c03u)), P1 = as_pair(__builtin_amdgcn_perm(b, a, 0x0c050c02u));
    const s16x2 P2 = as_pair(__builtin_amdgcn_perm(b, a, 0x0c060c01u)), P3 = as_pair(__builtin_amdgcn_perm(b, a, 0x0c070c00u));
    const int p0 = P0.x, q0 = P0.y, p1 = P1.x, q1 = P1.y, p2 = P2.x, q2 = P2.y;
    const int d0 = absdiff(p0, q0);
    const bool filt = on & (d0 < alpha) & (absdiff(p1, p0) < beta) & (absdiff(q1, q0) < beta);
    if (!any_lane(filt)) return;
    // ap / aq < beta (8-460, 8-461), only ever used for luma lines
    const bool apl = luma & (absdiff(p2, p0) < beta), aql = luma & (absdiff(q2, q0) < beta);
    s16x2 N0 = P0, N1 = P1, N2 = P2;
    const bool fn = filt & (bS < 4), fs = filt & (bS >= 4);
    const s16x2 X0 = pair_swap(P0), X1 = pair_swap(P1);  // (q0, p0), (q1, p1)
    if (any_lane(fn)) {  // bS < 4 (8.7.2.3)
        const int tc = tc0 + (int)apl + (int)aql + (int)!luma;  // luma tC0 + [ap] + [aq], chroma tC0 + 1
        const int delta = clip3(-tc, tc, (((q0 - p0) << 2) + (p1 - q1) + 4) >> 3);
        const short avg = (short)((p0 + q0 + 1) >> 1);
        // p'0 = Clip1(p0 + delta), q'0 = Clip1(q0 - delta)
        const s16x2 f0 = pair_clip(P0 + pair(delta, -delta), pair(0, 0), pair(255, 255));
        // p'1 = p1 + Clip3(-tC0, tC0, (p2 + avg - 2 p1) >> 1), q'1 alike, kept to 8 bits
        const s16x2 t = (P2 + (s16x2){avg, avg} - (P1 << 1)) >> 1;
        const s16x2 f1 = (P1 + pair_clip(t, pair(-tc0, -tc0), pair(tc0, tc0))) & (s16x2){255, 255};
        N0 = pair_sel(fn, fn, f0, N0);
        N1 = pair_sel(fn & apl, fn & aql, f1, N1);
    }
    if (any_lane(fs)) {  // bS == 4 (8.7.2.4)
        const bool strong = d0 < ((alpha >> 2) + 2);
        const bool sp = fs & apl & strong, sq = fs & aql & strong;
        const s16x2 m = P1 + P0 + X0;                                  // p1 + p0 + q0
        const s16x2 S0 = ((m << 1) + P2 + X1 + (s16x2){4, 4}) >> 3;    // p2 + 2p1 + 2p0 + 2q0 + q1 + 4
        const s16x2 W0 = ((P1 << 1) + P0 + X1 + (s16x2){2, 2}) >> 2;   // 2p1 + p0 + q1 + 2
        const s16x2 S1 = (m + P2 + (s16x2){2, 2}) >> 2;                // p2 + p1 + p0 + q0 + 2
        const s16x2 S2 = ((P3 << 1) + P2 * (s16x2){3, 3} + m + (s16x2){4, 4}) >> 3;  // 2p3 + 3p2 + p1 + p0 + q0 + 4
        N0 = pair_sel(sp, sq, S0, pair_sel(fs, fs, W0, N0));
        N1 = pair_sel(sp, sq, S1, N1);
        N2 = pair_sel(sp, sq, S2, N2);
    }
    // back to bytes: t01 = np0 np1 nq0 nq1, u = np0 np1 np2 nq2
    const uint32_t t01 = __builtin_amdgcn_perm(as_bits(N1), as_bits(N0), 0x06020400u);
    const uint32_t u = __builtin_amdgcn_perm(as_bits(N2), t01, 0x06040100u);
    w[J] = __builtin_amdgcn_perm(u, a, 0x04050600u);                                                  // p3 np2 np1 np0
    w[J + 1] = __builtin_amdgcn_perm(b, __builtin_amdgcn_perm(u, t01, 0x0c070302u), 0x07020100u);   // nq0 nq1 nq2 q3
}

// The edges of one direction (DIR 0: vertical edges, 1: horizontal) of the MB in the
// tile.  Lanes 0-15 take luma, 16-23 Cb, 24-31 Cr; lane k of a plane owns line k of the
// MB (a row for vertical edges, a column for horizontal ones) with the 4 samples before
// it, keeps the line in registers and runs its edges in order: left / top MB edge (when
// `edge`), then the internal ones (luma 4 / 12 skipped with the 8x8 transform, chroma
// only 4).  Lines of one direction touch disjoint samples, so this gives the reference's
// edge-by-edge order (8.7, Deblock.cpp:76-635).  qq / eq: deblocking words (see k_bs) of
// the MB and of the left / top MB.
// cur: tile column of the MB's luma column 0 (16 or 0; chroma cur / 2); the left MB sits
// in the other half of the tile, so a row's columns -4..-1 are at (16 - cur) + 12.
// PAIRED: lanes 32-63 do the same for a second MB on a tile of their own (the caller's
// per-lane wd / bsrec); act: the lane's MB is filtered in this step.
template <int DIR, bool PAIRED = false>
__device__ __forceinline__ void dbk_lines(const DevPic &P, uint32_t eq, bool edge, uint32_t qq,
                                          const uint8_t *bsrec, WaveDbk *wd, const DbkTables *tb, int cur,
                                          bool act = true) {
    const int lane = PAIRED ? (wave_lane() & 31) : wave_lane();
    if (!PAIRED && lane >= 32) return;
    const int plane = lane < 16 ? 0 : (lane < 24 ? 1 : 2);
    const bool luma = plane == 0;
    const int k = luma ? lane : (lane & 7);
    // tile row 0 of the plane; the MB's columns start at mc, the left MB's columns -4..-1 at lc
    uint8_t *base = luma ? &wd->L[0][0] : &wd->C[plane - 1][0][0];
    const int stride = luma ? 32 : 16, mc = luma ? cur : cur >> 1, lc = luma ? 28 - cur : 12 - (cur >> 1);
    uint32_t w[5];
    if (DIR == 0) {
        uint8_t *row = base + (k + 4) * stride;
        w[0] = *(const uint32_t *)(row + lc);
        w[1] = *(const uint32_t *)(row + mc);
        w[2] = *(const uint32_t *)(row + mc + 4);
        w[3] = luma ? *(const uint32_t *)(row + mc + 8) : 0u;
        w[4] = luma ? *(const uint32_t *)(row + mc + 12) : 0u;
    } else {
        // the quad of lanes that owns columns 4m..4m+3 reads rows 4i..4i+3 of them as dwords
        // (its lane j row 4i + j) and transposes them into the four columns; quads never
        // mix planes, so a quad agrees on `luma`
        const uint8_t *c = base + mc + (k & ~3) + (k & 3) * stride;
#pragma unroll
        for (int i = 0; i < 5; i++)
            w[i] = quad_transpose_bytes((i < 3 || luma) ? *(const uint32_t *)(c + 4 * i * stride) : 0u, k & 3);
    }
    const bool own = qq & (DIR == 0 ? DBQ_OWN_LEFT : DBQ_OWN_TOP);
    const int qpq = dbq_qp(qq, plane), qpe = edge && !own ? dbq_qp(eq, plane) : qpq;
    const int qav = (qpe + qpq + 1) >> 1;  // internal edges: (qpq + qpq + 1) >> 1 = qpq
    const int iA0 = clip3(0, 51, qav + P.offA), iB0 = clip3(0, 51, qav + P.offB);
    const int iAi = clip3(0, 51, qpq + P.offA), iBi = clip3(0, 51, qpq + P.offB);
    const bool t8 = dbq_t8(qq), on = act && P.dbk_idc != 1;
    const int seg = luma ? (k >> 2) : (k >> 1);  // 4-line luma segment of the line
    // bS of this segment's edges 0..3 (k_bs stores them [direction][segment][edge]); chroma
    // edge j lies on luma edge 2j.  All table lookups are issued before the filtering.
    const uint32_t bsv = *(const uint32_t *)&bsrec[DIR * 16 + seg * 4];
    const int bs0 = bsv & 255u, bs1 = luma ? (bsv >> 8) & 255u : (bsv >> 16) & 255u, bs2 = (bsv >> 16) & 255u,
              bs3 = bsv >> 24;
    int a0 = tb->alpha[iA0], b0 = tb->beta[iB0], ai = tb->alpha[iAi], bi = tb->beta[iBi];
    int t0 = tb->tc0[bs0 >= 1 && bs0 <= 3 ? bs0 - 1 : 0][iA0], t1 = tb->tc0[bs1 >= 1 && bs1 <= 3 ? bs1 - 1 : 0][iAi];
    int t2 = tb->tc0[bs2 >= 1 && bs2 <= 3 ? bs2 - 1 : 0][iAi], t3 = tb->tc0[bs3 >= 1 && bs3 <= 3 ? bs3 - 1 : 0][iAi];
    // one LDS round trip for the eight lookups: without this the compiler moves each into
    // the edge that first uses it and waits on it there
    asm volatile("" : "+v"(a0), "+v"(b0), "+v"(ai), "+v"(bi), "+v"(t0), "+v"(t1), "+v"(t2), "+v"(t3));
    dbk_line_edge<0>(w, on && edge, luma, bs0, a0, b0, t0);
    dbk_line_edge<1>(w, on && (luma ? !t8 : true), luma, bs1, ai, bi, t1);
    dbk_line_edge<2>(w, on && luma, luma, bs2, ai, bi, t2);
    dbk_line_edge<3>(w, on && luma && !t8, luma, bs3, ai, bi, t3);
    if (DIR == 0) {
        uint8_t *row = base + (k + 4) * stride;
        *(uint32_t *)(row + lc) = w[0];
        *(uint32_t *)(row + mc) = w[1];
        *(uint32_t *)(row + mc + 4) = w[2];
        if (luma) {
            *(uint32_t *)(row + mc + 8) = w[3];
            *(uint32_t *)(row + mc + 12) = w[4];
        }
    } else {
        // back to rows (the first and last row, p3 / q3 of the outer edges, keep their values)
        uint8_t *c = base + mc + (k & ~3) + (k & 3) * stride;
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const uint32_t v = quad_transpose_bytes(w[i], k & 3);
            if (i < 3 || luma) *(uint32_t *)(c + 4 * i * stride) = v;
        }
    }
}

// The MB's vertical edges, then its horizontal edges, on the tile (8.7, Deblock.cpp:76-635).
// lq / tq are the deblocking words of the left / top MB (read only when present).  bsrec: the MB's 32
// boundary strengths (device memory or an LDS copy).  The caller has synchronised the
// tile; it is complete when this returns.
__device__ __forceinline__ void dbk_filter(const DevPic &P, uint32_t qq, uint32_t lq, uint32_t tq, bool left, bool top,
                           const uint8_t *bsrec, WaveDbk *wd, const DbkTables *tb) {
    dbk_lines<0>(P, lq, left, qq, bsrec, wd, tb, DT_LX);
    wave_sync();
    dbk_lines<1>(P, tq, top, qq, bsrec, wd, tb, DT_LX);
    wave_sync();
}

// Writes back the MB and the strips its MB edges modified (3 columns / rows).
__device__ void dbk_store(const DbkPlanes &d, int mbx, int mby, bool left, bool top, const WaveDbk *wd) {
    const int lane = wave_lane();
    for (int i = lane; i < 400; i += WAVE) {
        int y = i / 20 - 4, x = i % 20 - 4;
        if (x < -3 || y < -3 || (x < 0 && y < 0) || (x < 0 && !left) || (y < 0 && !top)) continue;
        d.y[(uint64_t)(mby * 16 + y) * d.wl + mbx * 16 + x] = wd->L[y + 4][x + DT_LX];
    }
    for (int i = lane; i < 288; i += WAVE) {
        int c = i / 144, k = i % 144, y = k / 12 - 4, x = k % 12 - 4;
        if (x < -3 || y < -3 || (x < 0 && y < 0) || (x < 0 && !left) || (y < 0 && !top)) continue;
        uint8_t *pl = c == 0 ? d.cb : d.cr;
        pl[(uint64_t)(mby * 8 + y) * d.wc + mbx * 8 + x] = wd->C[c][y + 4][x + DT_CX];
    }
}

// One MB with every neighbour sample loaded from the picture (per-MB wavefront).
__device__ void deblock_mb(const DevPic &P, int addr, WaveDbk *wd, const DbkTables *tb) {
    const int W = P.W;
    const int mbx = addr % W, mby = addr / W;
    const DbkPlanes d = dbk_planes(P);
    const bool left = mbx > 0, top = mby > 0;
    dbk_load(d, mbx, mby, 1 | (top ? 2 : 0) | (left ? 4 : 0), wd);
    const uint32_t qq = G(P.dbq)[addr], lq = left ? G(P.dbq)[addr - 1] : 0u, tq = top ? G(P.dbq)[addr - W] : 0u;
    wave_sync();
    dbk_filter(P, qq, lq, tq, left, top, G(P.bs) + (size_t)addr * 32, wd, tb);
    dbk_store(d, mbx, mby, left, top, wd);
}

// After MB x of a row: its columns 12..15 become the left columns of MB x + 1.
__device__ void dbk_carry_left(WaveDbk *wd) {
    const int lane = wave_lane();
    wave_sync();
    for (int i = lane; i < 16 * 4 + 2 * 8 * 4; i += WAVE) {
        if (i < 64) {
            const int y = i >> 2, x = i & 3;
            wd->L[y + 4][DT_LX - 4 + x] = wd->L[y + 4][DT_LX + 12 + x];
        } else {
            const int k = i - 64, c = k >> 5, y = (k >> 2) & 7, x = k & 3;
            wd->C[c][y + 4][DT_CX - 4 + x] = wd->C[c][y + 4][DT_CX + 4 + x];
        }
    }
    wave_sync();
}

__device__ __forceinline__ uint64_t ld_acquire64(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until row progress word *w >= want (epoch-tagged: epoch << 32 | MBs done).
__device__ void wait_row(const uint64_t *w, uint64_t want, uint32_t *err) {
    for (uint32_t it = 0;; ++it) {
        const uint64_t v = ld_acquire64(w);
        const uint32_t lo = uniform((uint32_t)v), hi = uniform((uint32_t)(v >> 32));
        if ((((uint64_t)hi << 32) | lo) >= want) break;
        if (it >= SPIN_LIMIT) {
            atomicOr(err, DEVERR_TIMEOUT);
            break;
        }
        if ((it & 255u) == 255u && uniform(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0u) break;
        __builtin_amdgcn_s_sleep(1);
    }
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
}

// Row progress across CUs: epoch-tagged 64-bit words in global memory (agent scope).
struct RowSyncGlobal {
    GAS uint64_t *prog;
    uint64_t tag;
    uint32_t *err;
    __device__ void wait(int row, int need) { wait_row(&prog[row], tag | (uint32_t)need, err); }
    __device__ void post(int row, int done) {
        wave_sync();  // every lane's stores before the (wave-wide) release store
        __hip_atomic_store(&prog[row], tag | (uint32_t)done, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// Row progress of one picture inside one workgroup: counters in LDS, workgroup scope.
// The waves share the CU's caches, so no L2 write-back or invalidation is involved.
constexpr uint32_t LDS_SPIN_LIMIT = 1u << 22;
struct RowSyncLds {
    uint32_t *prog;
    uint32_t *err;
    __device__ void wait(int row, int need) {
        for (uint32_t it = 0;; ++it) {
            const uint32_t v = uniform(__hip_atomic_load(&prog[row], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
            if ((int)v >= need) break;
            if (it >= LDS_SPIN_LIMIT) {
                atomicOr(err, DEVERR_TIMEOUT);
                break;
            }
            if ((it & 1023u) == 1023u && uniform(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0u) break;
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    __device__ void post(int row, int done) {
        wave_sync();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __hip_atomic_store(&prog[row], (uint32_t)done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
};

// One whole MB row, left to right, in raster order.  MB (x, y) needs MB (x + 1, y - 1)
// done (its left-edge filter changed columns 13..15 of (x, y - 1), which (x, y)'s top
// edge reads), i.e. the row above at progress min(x + 2, W).  The left neighbour's
// columns stay in LDS from the previous step.
template <class Sync>
__device__ void deblock_row(const DevPic &P, int mby, Sync &sync, WaveDbk *wd, const DbkTables *tb) {
    const int W = P.W;
    const DbkPlanes d = dbk_planes(P);
    const bool top = mby > 0;
    uint32_t lq = 0u, qq = 0u, tq = 0u;
    for (int mbx = 0; mbx < W; mbx++) {
        const int addr = mby * W + mbx;
        const bool left = mbx > 0;
        dbk_load(d, mbx, mby, 1, wd);  // the MB's own samples: final before this launch
        lq = qq;
        qq = G(P.dbq)[addr];
        if (top) {
            tq = G(P.dbq)[addr - W];
            sync.wait(mby - 1, min(mbx + 2, W));
            dbk_load(d, mbx, mby, 2, wd);
        }
        wave_sync();
        dbk_filter(P, qq, lq, tq, left, top, G(P.bs) + (size_t)addr * 32, wd, tb);
        dbk_store(d, mbx, mby, left, top, wd);
        sync.post(mby, mbx + 1);
        dbk_carry_left(wd);
    }
}

// ------------------------------------------------------------ kernels
constexpr int DBK_WG_MAX_WAVES = 12;  // deblocking bands: up to 24 MB rows
// The LDS working sets are per wave and the per-MB code synchronises with wave barriers,
// so a workgroup may hold one wave (most kernels) or several (the band kernels).
// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b on XCD b mod 8), each
// XCD with its own L2.  Item of workgroup b (of n, one item each) such that every XCD
// takes one contiguous run of items: neighbouring MBs then share that XCD's L2 for
// their overlapping reference lines and for the cache lines their rows share.
constexpr uint32_t N_XCD = 8;
__device__ __forceinline__ uint32_t xcd_item(uint32_t b, uint32_t n) {
    const uint32_t x = b % N_XCD, q = n / N_XCD, r = n % N_XCD;
    return x * q + min(x, r) + b / N_XCD;
}

// Workgroups 0 .. n - 1 reconstruct the inter (and I_PCM) MBs of `items`, the following
// ni build the residuals of the intra MBs of `iitems`.
// MBAFF: the instance for MBAFF frames (field MBs: geometry, field scans, field references);
// frame pictures run the instance without any of it.
template <bool MBAFF, bool SCALED>
__global__ __launch_bounds__(WAVE) void k_inter(const DevPic *__restrict__ pics, const uint32_t *__restrict__ items, uint32_t n,
                                                const uint32_t *__restrict__ iitems, uint32_t ni, uint32_t *err) {
    __shared__ WaveRes lds;
    const uint32_t total = n + ni;
    if (blockIdx.x >= total) return;
    const uint32_t i = xcd_item(blockIdx.x, total);
    const uint32_t it = i < n ? items[i] : iitems[i - n];
    const DevPic &P = pics[it >> WI_PIC_SHIFT];
    if (i < n) inter_mb<MBAFF, SCALED>(P, (int)(it & WI_MB_MASK), &lds);
    else intra_res_build<MBAFF, SCALED>(P, (int)(it & WI_MB_MASK), &lds);
    (void)err;
}

// Neighbour availability of an intra MB for intra prediction (6.4.11, with
// constrained_intra_pred_flag); deps[k] = address of neighbour k (A, B, C, D) when it is
// itself intra-predicted in this picture (its samples come from the intra stage) and its
// samples are read, else -1.
__device__ __forceinline__ Avail intra_avail(const DevPic &P, int addr, int deps[4]) {
    const int W = P.W, mbx = addr % W, mby = addr / W;
    Avail av;
    av.m = (mbx > 0 ? 1u : 0u) | (mby > 0 ? 2u : 0u) | (mby > 0 && mbx < W - 1 ? 4u : 0u) | (mbx > 0 && mby > 0 ? 8u : 0u);
    const int nb[4] = {addr - 1, addr - W, addr - W + 1, addr - W - 1};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        deps[k] = -1;
        if (!(av.m & (1u << k))) continue;
        const int c = G(P.mbs)[nb[k]].mb_type;
        if (P.multi_slice && G(P.mbs)[nb[k]].slice_num != G(P.mbs)[addr].slice_num) {  // another slice (6.4.8)
            av.m &= ~(1u << k);
            continue;
        }
        if (P.cip && cip_hides(G(P.mbs)[nb[k]])) av.m &= ~(1u << k);
        if (cls_intra(c)) deps[k] = nb[k];
    }
    // Samples of C (above right) are read only by Intra_8x8 (the top reference samples
    // p[8..15, -1] of block 1 and their filtering, 8.3.2.2) and by Intra_4x4 block 5 in
    // the modes that use p[4..7, -1] (Diagonal_Down_Left 3, Vertical_Left 7); without
    // them the MB does not wait for C.
    const h264g_mb &m = G(P.mbs)[addr];
    const bool uses_c = m.mb_type == H264G_MB_I8x8 ||
                        (m.mb_type == H264G_MB_I4x4 && (m.intra_pred_modes[5] == 3 || m.intra_pred_modes[5] == 7));
    if (!uses_c) deps[2] = -1;
    return av;
}

// The same for an MB of an MBAFF frame (6.4.10 / Table 6-4; no constrained intra):
// availability of the samples left (A), above (B), above right (C) and above left (D) of
// the MB's view.  A frame MB at the bottom of its pair reads its top MB above it and the
// left pair above left, and has no C; every other MB reads the pairs around its own.  It
// waits for the intra MBs of the pairs whose samples it may read (both MBs of a pair: the
// rows of a frame MB beside a field pair come from both, and the other way round), and
// for the pair above right only when it reads C; deps holds up to 8 addresses.
__device__ __forceinline__ Avail intra_avail_mbaff(const DevPic &P, int addr, int deps[8]) {
    const int W = P.W, pair = addr >> 1, px = pair % W, py = pair / W, bot = addr & 1;
    const bool field = (G(P.mbs)[addr].flags & H264G_MBF_FIELD) != 0, in_pair = !field && bot;
    const bool A = px > 0, B = py > 0, C = py > 0 && px < W - 1, D = py > 0 && px > 0;
    Avail av;
    av.m = (A ? 1u : 0u) | ((in_pair || B) ? 2u : 0u) | (!in_pair && C ? 4u : 0u) | ((in_pair ? A : D) ? 8u : 0u);
    const h264g_mb &m = G(P.mbs)[addr];
    const bool uses_c = m.mb_type == H264G_MB_I8x8 ||
                        (m.mb_type == H264G_MB_I4x4 && (m.intra_pred_modes[5] == 3 || m.intra_pred_modes[5] == 7));
    // neighbouring pairs read: left; above, above-left, above-right (not for the bottom
    // frame MB, whose row above is its own top MB, an earlier ticket of this launch)
    const int pn[4] = {A ? pair - 1 : -1, !in_pair && B ? pair - W : -1, !in_pair && D ? pair - W - 1 : -1,
                       !in_pair && C && uses_c ? pair - W + 1 : -1};
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int a = pn[k] < 0 ? -1 : 2 * pn[k] + h;
            deps[2 * k + h] = a >= 0 && cls_intra(G(P.mbs)[a].mb_type) ? a : -1;
        }
    if (in_pair && cls_intra(G(P.mbs)[addr - 1].mb_type)) deps[2] = addr - 1;  // slot of the unused above pair
    return av;
}

template <bool MBAFF>
__global__ __launch_bounds__(WAVE) void k_intra(const DevPic *__restrict__ pics, const uint32_t *__restrict__ items, uint32_t n,
                                                DevCtl *ctl, uint32_t *epoch_p, uint32_t *err) {
    __shared__ WaveIntra lds;
    const uint32_t epoch = launch_epoch(epoch_p);
    for (;;) {
        int t = next_ticket(ctl, n);
        if (t < 0) break;
        const uint32_t it = items[t];
        const DevPic &P = pics[it >> WI_PIC_SHIFT];
        const int addr = (int)(it & WI_MB_MASK);
        int deps[MBAFF ? 8 : 4];
        const Avail av = MBAFF ? intra_avail_mbaff(P, addr, deps) : intra_avail(P, addr, deps);
        const IntraRes r = intra_res_fetch(P, addr);
        if (G(P.mbs)[addr].mb_type == H264G_MB_I4x4) intra4_prep(&G(P.mbs)[addr], av, &lds);
        else if (G(P.mbs)[addr].mb_type == H264G_MB_I8x8) intra8_prep(&G(P.mbs)[addr], av, &lds);
        wait_deps(G(P.flags), deps, MBAFF ? 8 : 4, epoch, err);
        intra_res_place(r, &lds);
        intra_mb<MBAFF>(P, addr, &lds, av, r.res);
        publish(&G(P.flags)[addr], epoch);
    }
    retire(ctl, gridDim.x, epoch_p, epoch);
}

// ---- bands of MB rows, one workgroup per band
// A band is consecutive MB rows of one picture, one row per wave (intra) or two
// (deblocking, see deblock_pair_band).  Inside the band the rows synchronise through LDS
// counters (workgroup scope: the waves share the CU's caches).  A band's last row hands
// what the next band's first row reads (in another workgroup, possibly on another XCD)
// through device memory as epoch-tagged 64-bit words, one agent-scope atomic each
// (strip_put_g, istrip_put): the reader polls the words themselves, and no agent-scope
// release / acquire (an L2 write-back or invalidation of the whole XCD) is issued.  Work
// items are (picture, band) in band-major order and a workgroup takes its item by ticket
// when it starts, so the band above always belongs to a workgroup that started earlier
// and never waits on a later one.

// Takes the workgroup's ticket (into *slot) and clears `n` LDS counters.
__device__ __forceinline__ uint32_t band_ticket(DevCtl *ctl, uint32_t *slot, uint32_t *ctr, int n) {
    if (threadIdx.x == 0) *slot = atomicAdd(&ctl->ticket, 1u);
    if ((int)threadIdx.x < n) ctr[threadIdx.x] = 0;
    __syncthreads();
    return uniform(*slot);
}
// Every wave of the workgroup is done: the launch's last workgroup resets the control
// block and advances the epoch (see retire).
__device__ __forceinline__ void band_retire(DevCtl *ctl, uint32_t *epoch, uint32_t cur) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t d = atomicAdd(&ctl->done, 1u);
        if (d == gridDim.x - 1) {
            __hip_atomic_store(&ctl->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(epoch, cur + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}
// Intra strips: the bottom row of each intra MB of a band's last MB row (luma 0..15, Cb
// 0..7, Cr 0..7) as 8 words of 64 bits, 4 samples and the launch epoch each, for the
// next band's first row.  As with the deblocking strips (strip_put_g), each word is one
// agent-scope atomic and the reader waits until the words carry its epoch, so the
// samples cross workgroups (and XCDs) without a fence; the picture copy of those samples
// is not read across bands.
__device__ __forceinline__ void istrip_put(uint64_t *sw, WaveIntra *wi, uint32_t epoch) {
    const int lane = wave_lane();
    wave_sync();
    if (lane < 8) {
        const uint8_t *src = lane < 4 ? &wi->T[16][1 + 4 * lane] : &wi->cbot[(lane - 4) >> 1][4 * (lane & 1)];
        const uint32_t v = (uint32_t)src[0] | (uint32_t)src[1] << 8 | (uint32_t)src[2] << 16 | (uint32_t)src[3] << 24;
        __hip_atomic_store(&sw[lane], (uint64_t)epoch << 32 | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
// The row -1 samples of the intra neighbours in `above` (2: B = MB x, 4: C = x + 1, 8: D =
// x - 1) from the previous band's strips (sw: its row of W x 8 words) into the tile.
// Lanes 0..23: MB x - 1 + (lane >> 3), word lane & 7 (D: the last sample of each plane,
// words 3, 5, 7; B: all; C: luma 0..7, words 0, 1).  Bounded like the other waits.
__device__ void istrip_get(const uint64_t *sw, int x, uint32_t above, WaveIntra *wi, uint32_t epoch, uint32_t *err) {
    const int lane = wave_lane(), m = lane >> 3, k = lane & 7;
    const uint32_t bit = m == 0 ? 8u : (m == 1 ? 2u : 4u);
    const bool mine = lane < 24 && (above & bit) && (m == 1 || (m == 0 ? (k == 3 || k == 5 || k == 7) : k < 2));
    uint64_t v = (uint64_t)epoch << 32;
    for (uint32_t it = 0;; ++it) {
        if (mine) v = __hip_atomic_load(&sw[(size_t)(x - 1 + m) * 8 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__ballot(mine && (uint32_t)(v >> 32) != epoch) == 0) break;
        if (it >= SPIN_LIMIT) {
            atomicOr(err, DEVERR_TIMEOUT);
            break;
        }
        if ((it & 255u) == 255u && uniform(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0u) break;
        __builtin_amdgcn_s_sleep(1);
    }
    if (mine) {
        const uint32_t d = (uint32_t)v;
        if (m == 0) {
            const uint8_t s = (uint8_t)(d >> 24);
            if (k == 3) wi->T[0][0] = s;
            else wi->C[(k - 5) >> 1][0][0] = s;
        } else {
            uint8_t *dst = m == 2 ? &wi->T[0][17 + 4 * k] : (k < 4 ? &wi->T[0][1 + 4 * k] : &wi->C[(k - 4) >> 1][0][1 + 4 * (k & 1)]);
#pragma unroll
            for (int t = 0; t < 4; t++) dst[t] = (uint8_t)(d >> (8 * t));
        }
    }
}

constexpr int INTRA_WG_MAX_WAVES = 8;
__global__ __launch_bounds__(WAVE * INTRA_WG_MAX_WAVES) void k_intra_band(const DevPic *__restrict__ pics,
                                                                          const uint32_t *__restrict__ items, uint32_t n,
                                                                          DevCtl *ctl, uint32_t *epoch_p, uint32_t *err) {
    __shared__ WaveIntra tiles[INTRA_WG_MAX_WAVES];
    __shared__ uint32_t prog[INTRA_WG_MAX_WAVES];
    __shared__ uint32_t s_ticket;
    const uint32_t epoch = launch_epoch(epoch_p);
    const int nw = blockDim.x / WAVE, w = (int)uniform(threadIdx.x / WAVE);
    const uint32_t t = band_ticket(ctl, &s_ticket, prog, nw);
    if (t < n) {
        const uint32_t it = items[t];
        const DevPic &P = pics[it >> WI_PIC_SHIFT];
        const int band = (int)(it & WI_MB_MASK);
        const int W = P.W, H = P.H, y0 = band * nw, y = y0 + w, y1 = min(H, y0 + nw);
        if (P.n_intra != 0 && y < H) {
            const bool from_g = y > 0 && y == y0, to_g = y < H - 1 && y == y1 - 1;
            uint64_t *iout = G(P.istrip) + (size_t)band * W * 8;                 // this band -> next
            const uint64_t *iin = G(P.istrip) + (size_t)(band - 1) * W * 8;      // previous band -> this
            RowSyncLds sync{prog, err};
            const uint32_t *off = G(P.irow), *col = G(P.irow) + H + 1;
            const uint32_t b = uniform(off[y]), e = uniform(off[y + 1]);
            auto post = [&](uint32_t done) {  // progress for the next row of the band
                if (!to_g && y < H - 1) sync.post(y - y0, (int)done);
            };
            post(b < e ? uniform(col[b]) & ICOL_X_MASK : (uint32_t)W);
            for (uint32_t i = b; i < e; i++) {
                // the MB's availability and intra neighbours come with its entry, so that
                // nothing but its own record and residual is loaded before the wait
                const uint32_t ent = uniform(col[i]);
                const int x = (int)(ent & ICOL_X_MASK), addr = y * W + x;
                Avail av;
                av.m = (ent >> ICOL_AV_SHIFT) & 15u;
                const uint32_t dep = (ent >> ICOL_DEP_SHIFT) & 15u;
                IntraRes r;  // in flight during the wait
                r.res = (ent & ICOL_RES) != 0;
                r.v = make_uint4(0, 0, 0, 0);
                if (r.res && wave_lane() < 48) r.v = ((const uint4 *)(G(P.ires) + (size_t)i * 384))[wave_lane()];
                const h264g_mb *mb = &G(P.mbs)[addr];
                if (mb->mb_type == H264G_MB_I4x4) intra4_prep(mb, av, &tiles[w]);
                else if (mb->mb_type == H264G_MB_I8x8) intra8_prep(mb, av, &tiles[w]);
                // the row above up to C (x + 1) or up to B (x) when only B / D are read and intra
                const int need = (dep & 4u) ? min(x + 2, W) : ((dep & 10u) ? x + 1 : 0);
                uint32_t above = 0;
                if (need > 0) {
                    if (from_g) {
                        above = dep & 14u;
                        istrip_get(iin, x, above, &tiles[w], epoch, err);
                    } else {
                        sync.wait(y - 1 - y0, need);
                    }
                }
                intra_res_place(r, &tiles[w]);
                intra_mb<false>(P, addr, &tiles[w], av, r.res, above);
                if (to_g) istrip_put(iout + (size_t)x * 8, &tiles[w], epoch);
                const bool end = i + 1 == e;
                post(end ? (uint32_t)W : uniform(col[i + 1]) & ICOL_X_MASK);
            }
        }
    }
    band_retire(ctl, epoch_p, epoch);
}

__global__ __launch_bounds__(WAVE) void k_deblock(const DevPic *__restrict__ pics, const uint32_t *__restrict__ items, uint32_t n,
                                                  DevCtl *ctl, uint32_t *epoch_p, uint32_t *err) {
    __shared__ WaveDbk lds;
    __shared__ DbkTables tb;
    dbk_tables_init(&tb);
    __syncthreads();
    const uint32_t epoch = launch_epoch(epoch_p);
    for (;;) {
        int t = next_ticket(ctl, n);
        if (t < 0) break;
        const uint32_t it = items[t];
        const DevPic &P = pics[it >> WI_PIC_SHIFT];
        const int addr = (int)(it & WI_MB_MASK), W = P.W, mbx = addr % W, mby = addr / W;
        int deps[2] = {mbx > 0 ? addr - 1 : -1, mby > 0 ? (mbx < W - 1 ? addr - W + 1 : addr - W) : -1};
        wait_deps(G(P.flags), deps, 2, epoch, err);
        deblock_mb(P, addr, &lds, &tb);
        publish(&G(P.flags)[addr], epoch);
    }
    retire(ctl, gridDim.x, epoch_p, epoch);
}

// ---- deblocking of MBAFF frames, one MB pair per wave
// 8.7 in MB address order (Deblock.cpp:76-635): the pair's top MB, then its bottom MB.
// Pair (x, y) follows pair (x - 1, y), whose right columns its left edges filter, and
// pair (x + 1, y - 1), whose left edges filter the right columns of pair (x, y - 1) that
// its top edges read: persistent waves over pairs in t = x + 2 y order, one progress word
// per pair (G(P.pairprog)).  The wave works on an LDS copy of the samples its edges touch:
// luma rows -8..31 (a field MB's top edge reads four rows of its parity above the pair)
// by columns -4..15, chroma rows -8..15 by columns -4..7, written back afterwards.
constexpr int MR_LROWS = 40, MR_LCOLS = 20, MR_CROWS = 24, MR_CCOLS = 12, MR_ROW0 = 8, MR_COL0 = 4;
struct alignas(16) MbaffRegion {
    uint8_t L[MR_LROWS][MR_LCOLS];
    uint8_t C[2][MR_CROWS][MR_CCOLS];
};

// Copies the pair's region between the picture and LDS (to_lds: load, else store).
// Columns left of the picture and rows above it are skipped.
__device__ void mbaff_region(const DevPic &P, int px, int py, MbaffRegion *rg, bool to_lds) {
    const int lane = wave_lane(), wl = P.W * 16, wc = P.W * 8;
    uint8_t *y = slot_base(P, P.dst), *cb = y + (uint64_t)wl * (P.H * 16), *cr = cb + (uint64_t)wc * (P.H * 8);
    // dwords: luma rows x 5, then chroma planes x rows x 3
    constexpr int NL = MR_LROWS * 5, NC = MR_CROWS * 3;
    for (int i = lane; i < NL + 2 * NC; i += WAVE) {
        uint8_t *pic, *lds;
        int r, d;
        if (i < NL) {
            r = i / 5 - MR_ROW0;
            d = i % 5;
            if ((py == 0 && r < 0) || (px == 0 && d == 0)) continue;
            pic = y + (uint64_t)(py * 32 + r) * wl + px * 16 - 4 + 4 * d;
            lds = &rg->L[r + MR_ROW0][4 * d];
        } else {
            const int k = i - NL, c = k / NC;
            r = (k % NC) / 3 - MR_ROW0;
            d = (k % NC) % 3;
            if ((py == 0 && r < 0) || (px == 0 && d == 0)) continue;
            pic = (c ? cr : cb) + (uint64_t)(py * 16 + r) * wc + px * 8 - 4 + 4 * d;
            lds = &rg->C[c][r + MR_ROW0][4 * d];
        }
        if (to_lds) *(uint32_t *)lds = *(const uint32_t *)pic;
        else *(uint32_t *)pic = *(const uint32_t *)lds;
    }
}

// 8.7.2.1 (Deblock.cpp:994-1310) in an MBAFF frame, for the luma positions (xp, yp) of p0
// in MB ap and (xq, yq) of q0 in MB aq: mixedModeEdgeFlag between a frame and a field MB;
// bS 4 at MB edges between frame MBs and at vertical MB edges, 3 at the other intra edges;
// 2 for levels; 1 at mixed edges; then the motion tests with a vertical limit of 2
// quarter samples when q is a field MB (:1157-1158), comparing RefPicListX[RefIdxLX]
// (G(P.bsref), quirk Q4) with quirk Q3.  I_PCM counts as neither intra nor inter.
__device__ __forceinline__ int bs_mbaff(const DevPic &P, int ap, int aq, int xp, int yp, int xq, int yq, bool vertical) {
    const h264g_mb &mp = G(P.mbs)[ap], &mq = G(P.mbs)[aq];
    const DevAux xa = G(P.aux)[ap], xb = G(P.aux)[aq];
    const bool pf = mp.flags & H264G_MBF_FIELD, qf = mq.flags & H264G_MBF_FIELD;
    const bool intra = ((xa.kind | xb.kind) & AUX_INTRA) != 0, mixed = ap != aq && pf != qf;
    if (ap != aq && intra && ((!pf && !qf) || vertical)) return 4;
    if (intra && (!mixed || !vertical)) return 3;
    const int bp4 = 8 * (yp >> 3) + 4 * (xp >> 3) + 2 * ((yp & 7) >> 2) + ((xp & 7) >> 2);
    const int bq4 = 8 * (yq >> 3) + 4 * (xq >> 3) + 2 * ((yq & 7) >> 2) + ((xq & 7) >> 2);
    const int bp8 = 2 * (yp >> 3) + (xp >> 3), bq8 = 2 * (yq >> 3) + (xq >> 3);
    const bool pt8 = xa.kind & AUX_T8, qt8 = xb.kind & AUX_T8;
    if ((pt8 && ((xa.nz8 >> bp8) & 1)) || (!pt8 && ((xa.nz4 >> bp4) & 1)) || (qt8 && ((xb.nz8 >> bq8) & 1)) ||
        (!qt8 && ((xb.nz4 >> bq4) & 1)))
        return 2;
    if (mixed) return 1;
    const int ylim = qf ? 2 : 4;
    const int qp = bp8, qq = bq8, rp = (yp >> 2) * 4 + (xp >> 2), rq = (yq >> 2) * 4 + (xq >> 2);
    const int p0 = G(P.bsref)[ap * 8 + qp], p1 = G(P.bsref)[ap * 8 + 4 + qp], q0 = G(P.bsref)[aq * 8 + qq], q1 = G(P.bsref)[aq * 8 + 4 + qq];
    const int fp0 = (xa.pf >> qp) & 1, fp1 = (xa.pf >> (4 + qp)) & 1, fq0 = (xb.pf >> qq) & 1, fq1 = (xb.pf >> (4 + qq)) & 1;
    if (!(((p0 == q0 && p1 == q1) || (p0 == q1 && p1 == q0)) && fp0 + fp1 == fq0 + fq1)) return 1;
    auto far = [&](int ax, int ay, int bx, int by) { return iabs(ax - bx) >= 4 || iabs(ay - by) >= ylim; };
    const int mp0x = mp.mv[0][rp][0], mp0y = mp.mv[0][rp][1], mp1x = mp.mv[1][rp][0], mp1y = mp.mv[1][rp][1];
    const int mq0x = mq.mv[0][rq][0], mq0y = mq.mv[0][rq][1], mq1x = mq.mv[1][rq][0], mq1y = mq.mv[1][rq][1];
    if (fp0 && !fp1 && fq0 && !fq1 && far(mp0x, mp0y, mq0x, mq0y)) return 1;
    if (fp0 && !fp1 && !fq0 && fq1 && far(mp0x, mp0y, mq1x, mq1y)) return 1;
    if (!fp0 && fp1 && fq0 && !fq1 && far(mp1x, mp1y, mq0x, mq0y)) return 1;
    if (!fp0 && fp1 && !fq0 && fq1 && far(mp1x, mp1y, mq1x, mq1y)) return 1;
    if (fp0 && fp1 && p0 != p1 && fq0 && fq1 && ((q0 == p0 && q1 == p1) || (q0 == p1 && q1 == p0))) {
        if (q0 == p0 && (far(mp0x, mp0y, mq0x, mq0y) || far(mp1x, mp1y, mq1x, mq1y))) return 1;
        if (q0 == p1 && (far(mp1x, mp1y, mq0x, mq0y) || far(mp0x, mp0y, mq1x, mq1y))) return 1;
    }
    if (fp0 && fp1 && p0 == p1 && fq0 && fq1 && q0 == q1 && q0 == p0) {
        if (far(mp0x, mp0y, mq0x, mq0y) || (far(mp1x, mp1y, mq1x, mq1y) && far(mp0x, mp0y, mq1x, mq1y)) ||
            far(mp1x, mp1y, mq0x, mq0y))
            return 1;
    }
    return 0;
}

// One edge pass of MB addr (8.7.1, Deblock.cpp:639-990) on the pair's LDS region: lane k
// of a plane (luma lanes 0-15, Cb 16-23, Cr 24-31; act: the lane's plane takes part) owns
// line k.  e: the edge's x (vertical) or y (horizontal) inside the MB; lines of a field-
// mode pass are two rows apart (dy = 1 + fieldMode), and yE % 2 picks the bottom-field
// lines of a frame MB's top edge over a field pair.  pn: the p-MB of an MB edge; with
// leftMix (a frame MB beside a field pair) it alternates with the line's parity while the
// p0 position keeps the current MB's line number (quirk Q6).
__device__ __forceinline__ void mbaff_edge(const DevPic &P, MbaffRegion *rg, const DbkTables *tb, int addr, int lrow0, int crow0,
                           bool act, bool vertical, int e, int fieldMode, int pn, bool leftMix) {
    const int lane = wave_lane();
    const int plane = lane < 16 ? 0 : (lane < 24 ? 1 : 2), chroma = plane != 0;
    const int k = chroma ? (lane & 7) : (lane & 15), nE = chroma ? 8 : 16, dy = 1 + fieldMode, s = chroma ? 2 : 1;
    act = act && lane < 32;
    uint8_t *base = chroma ? &rg->C[plane - 1][0][0] : &rg->L[0][0];
    const int pitch = chroma ? MR_CCOLS : MR_LCOLS, row0 = MR_ROW0 + (chroma ? crow0 : lrow0);
    const int xE = vertical ? e : k, yE = vertical ? k : e;
    int qi[4], pi[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (vertical) {
            qi[i] = (row0 + dy * yE) * pitch + MR_COL0 + xE + i;
            pi[i] = (row0 + dy * yE) * pitch + MR_COL0 + xE - i - 1;
        } else {
            qi[i] = (row0 + dy * (yE + i) - (yE & 1)) * pitch + MR_COL0 + xE;
            pi[i] = (row0 + dy * (yE - i - 1) - (yE & 1)) * pitch + MR_COL0 + xE;
        }
    }
    uint32_t w[5] = {0, 0, 0, 0, 0};
    int ap = addr, xp, yp, xq, yq;
    if (vertical) {
        xp = xE - 1;
        if (xp < 0) {
            if (pn >= 0) ap = leftMix ? pn + (yE & 1) : pn;
            xp += nE;
        }
        yp = yE; xq = xE; yq = yE;
    } else {
        xp = xE;
        yp = (yE - 1) - (yE & 1);
        if (yp < 0) {
            if (pn >= 0) ap = pn;
            yp += nE;
        }
        xq = xE; yq = yE - (yE & 1);
    }
    int bS = 0, alpha = 0, beta = 0, tc0 = 0;
    if (act) {
        w[0] = pack4(base[pi[3]], base[pi[2]], base[pi[1]], base[pi[0]]);
        w[1] = pack4(base[qi[0]], base[qi[1]], base[qi[2]], base[qi[3]]);
        bS = bs_mbaff(P, ap, addr, xp * s, yp * s, xq * s, yq * s, vertical);
        const int qpp = G(P.aux)[ap].qdb[plane], qpq = G(P.aux)[addr].qdb[plane], qav = (qpp + qpq + 1) >> 1;
        const int ia = clip3(0, 51, qav + P.offA), ib = clip3(0, 51, qav + P.offB);
        alpha = tb->alpha[ia];
        beta = tb->beta[ib];
        tc0 = bS >= 1 && bS <= 3 ? tb->tc0[bS - 1][ia] : 0;
    }
    dbk_line_edge<0>(w, act, !chroma, bS, alpha, beta, tc0);
    if (act) {
        base[pi[0]] = (uint8_t)(w[0] >> 24);
        base[pi[1]] = (uint8_t)(w[0] >> 16);
        base[pi[2]] = (uint8_t)(w[0] >> 8);
        base[qi[0]] = (uint8_t)w[1];
        base[qi[1]] = (uint8_t)(w[1] >> 8);
        base[qi[2]] = (uint8_t)(w[1] >> 16);
    }
    wave_sync();
}

// The edges of one MB of the pair in the reference's order (Deblock.cpp:76-635): luma
// left, internal vertical, top (twice, in field mode, for a frame MB under a field pair:
// its top field lines against the pair's top MB, its bottom field lines against the bottom
// MB; chroma against the bottom MB both times, quirk Q5), internal horizontal; chroma in
// the same order, each chroma pass beside a luma pass (different planes).
__device__ __forceinline__ void deblock_mb_mbaff(const DevPic &P, int addr, MbaffRegion *rg, const DbkTables *tb) {
    const int lane = wave_lane(), W = P.W, pair = addr >> 1, px = pair % W, py = pair / W, bot = addr & 1;
    const bool field = (G(P.mbs)[addr].flags & H264G_MBF_FIELD) != 0, t8 = (G(P.aux)[addr].kind & AUX_T8) != 0;
    const bool lfield = px > 0 && (G(P.mbs)[2 * (pair - 1)].flags & H264G_MBF_FIELD);
    const bool afield = py > 0 && (G(P.mbs)[2 * (pair - W)].flags & H264G_MBF_FIELD);
    // 6.4.11.1 with Table 6-4 at (-1, 0) and (0, -1)
    const int mbA = px > 0 ? 2 * (pair - 1) + (field == lfield ? bot : 0) : -1;
    const int mbB = !field && bot ? addr - 1 : (py > 0 ? 2 * (pair - W) + (field && !bot && afield ? 0 : 1) : -1);
    const bool left = px > 0, upper = !(py == 0 && (field || !bot));
    const bool leftMix = !field && lfield, topMix = !bot && py > 0 && !field && afield;
    const int lrow0 = field ? bot : 16 * bot, crow0 = field ? bot : 8 * bot, fm = field ? 1 : 0;
    const bool luma = lane < 16, chroma = lane >= 16;
    mbaff_edge(P, rg, tb, addr, lrow0, crow0, left, true, 0, fm, mbA, leftMix);
    mbaff_edge(P, rg, tb, addr, lrow0, crow0, chroma || !t8, true, 4, fm, mbA, false);
    mbaff_edge(P, rg, tb, addr, lrow0, crow0, luma, true, 8, fm, mbA, false);
    mbaff_edge(P, rg, tb, addr, lrow0, crow0, luma && !t8, true, 12, fm, mbA, false);
    if (topMix) {
        mbaff_edge(P, rg, tb, addr, lrow0, crow0, true, false, 0, 1, luma ? mbB - 1 : mbB, false);
        mbaff_edge(P, rg, tb, addr, lrow0, crow0, true, false, 1, 1, mbB, false);
    } else {
        mbaff_edge(P, rg, tb, addr, lrow0, crow0, upper, false, 0, fm, mbB, false);
    }
    mbaff_edge(P, rg, tb, addr, lrow0, crow0, chroma || !t8, false, 4, fm, mbB, false);
    mbaff_edge(P, rg, tb, addr, lrow0, crow0, luma, false, 8, fm, mbB, false);
    mbaff_edge(P, rg, tb, addr, lrow0, crow0, luma && !t8, false, 12, fm, mbB, false);
}

__global__ __launch_bounds__(WAVE) void k_deblock_mbaff(const DevPic *__restrict__ pics, const uint32_t *__restrict__ items,
                                                        uint32_t n, DevCtl *ctl, uint32_t *epoch_p, uint32_t *err) {
    __shared__ MbaffRegion rg;
    __shared__ DbkTables tb;
    dbk_tables_init(&tb);
    __syncthreads();
    const uint32_t epoch = launch_epoch(epoch_p);
    for (;;) {
        int t = next_ticket(ctl, n);
        if (t < 0) break;
        const uint32_t it = items[t];
        const DevPic &P = pics[it >> WI_PIC_SHIFT];
        const int pair = (int)(it & WI_MB_MASK), W = P.W, px = pair % W, py = pair / W;
        wait_dep(G(P.pairprog), px > 0 ? pair - 1 : -1, epoch, err);
        wait_dep(G(P.pairprog), py > 0 ? (px < W - 1 ? pair - W + 1 : pair - W) : -1, epoch, err);
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
        mbaff_region(P, px, py, &rg, true);
        wave_sync();
        deblock_mb_mbaff(P, 2 * pair, &rg, &tb);
        deblock_mb_mbaff(P, 2 * pair + 1, &rg, &tb);
        mbaff_region(P, px, py, &rg, false);
        publish(&G(P.pairprog)[pair], epoch);
    }
    retire(ctl, gridDim.x, epoch_p, epoch);
}

// Row-pipelined deblocking: a ticket is one MB row of one picture (queue ordered by
// row index across pictures, so the row above always holds an earlier ticket).
__global__ __launch_bounds__(WAVE) void k_deblock_rows(const DevPic *__restrict__ pics, const uint32_t *__restrict__ items,
                                                       uint32_t n, DevCtl *ctl, uint32_t *epoch_p, uint32_t *err) {
    __shared__ WaveDbk lds;
    __shared__ DbkTables tb;
    dbk_tables_init(&tb);
    __syncthreads();
    const uint32_t epoch = launch_epoch(epoch_p);
    const uint64_t tag = (uint64_t)epoch << 32;
    for (;;) {
        int t = next_ticket(ctl, n);
        if (t < 0) break;
        const uint32_t it = items[t];
        const DevPic &P = pics[it >> WI_PIC_SHIFT];
        RowSyncGlobal sync{G(P.rowprog), tag, err};
        deblock_row(P, (int)(it & WI_MB_MASK), sync, &lds, &tb);
    }
    retire(ctl, gridDim.x, epoch_p, epoch);
}

// ---- deblocking over bands, two MB rows per wave
// A band of R MB rows runs on ceil(R / 2) waves; wave w takes rows y0 + 2w (the upper
// half: lanes 0-31) and y0 + 2w + 1 (the lower half: lanes 32-63).  The halves run the
// per-MB step on 32 lanes each (luma lines on half-lanes 0-15, Cb 16-23, Cr 24-31), each
// on a tile of its own, the lower half one MB behind the upper: one instruction stream
// filters two MBs, and every lane of the wave has a line to filter.
//
// Row y hands the bottom 4 luma / chroma rows of each MB (the p samples of the next
// row's top edge) to row y + 1 as a strip: inside a wave from the upper tile to the
// lower tile, from a wave's lower row to the next wave's upper row through a ring of
// STRIP_SLOTS LDS slots, and from a band's last row to the next band through a
// full-width strip array in device memory (G(P.gstrip), one row of W strips per band
// boundary).  No sample otherwise travels between rows: each sample is written to the
// picture by exactly one half, row y storing rows 0..12 (chroma 0..4) of its MBs and row
// y + 1 the rows 13..15 (chroma 5..7) after its top-edge filtering (the picture's last
// row stores all).  An MB is stored, and its strip handed down, one step after its own
// filtering, when its right neighbour's left edge has finished its columns 13..15 (the
// row's last MB one step later as well, with nothing to filter in that step).  A lower
// row writes strip x into ring slot x mod STRIP_SLOTS only after the next wave's upper
// row has read strip x - STRIP_SLOTS; every wave of a band is resident, so the reader is
// always running.
constexpr int STRIP_SLOTS = 8;
struct alignas(16) StripSlot {
    uint32_t w[32];           // strip dword of half-lane j (see strip_word)
};
struct DbkWgShared {
    WaveDbk tiles[2 * DBK_WG_MAX_WAVES];               // tile of row y0 + r
    StripSlot ring[DBK_WG_MAX_WAVES][STRIP_SLOTS];    // wave w's lower row -> wave w + 1's upper row
    alignas(16) uint8_t bsrec[2 * DBK_WG_MAX_WAVES][32];
    uint32_t prog[DBK_WG_MAX_WAVES];   // wave w's lower row: strips of MBs 0 .. prog - 1 complete
    uint32_t cons[DBK_WG_MAX_WAVES];   // wave w's lower row: strips of MBs 0 .. cons - 1 read by wave w + 1
    DbkTables tb;
};

__device__ __forceinline__ void lds_release() {
    wave_sync();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
}
__device__ void lds_wait(const uint32_t *w, uint32_t need, uint32_t *err) {
    for (uint32_t it = 0;; ++it) {
        if (uniform(__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) >= need) break;
        if (it >= LDS_SPIN_LIMIT) {
            atomicOr(err, DEVERR_TIMEOUT);
            break;
        }
        if ((it & 1023u) == 1023u && uniform(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0u) break;
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
__device__ __forceinline__ void lds_post(uint32_t *w, uint32_t v) {
    lds_release();
    __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int half_lane() { return threadIdx.x & 31; }

// Body of MB (mbx, mby) and its boundary strengths for one half (prefetch registers):
// half-lane j holds luma row j / 2, bytes 8 (j & 1) .. + 7, the chroma dword j of the
// two 8x8 planes and, for j < 8, boundary-strength dword j.
struct DbkPrefetch {
    uint2 luma;
    uint32_t chroma, bs;
};
__device__ __forceinline__ DbkPrefetch dbk_fetch(const DevPic &P, const DbkPlanes &d, int mbx, int mby, bool on) {
    const int hl = half_lane();
    DbkPrefetch f{make_uint2(0u, 0u), 0u, 0u};
    if (on) {
        f.luma = *(const uint2 *)&d.y[(uint64_t)(mby * 16 + (hl >> 1)) * d.wl + mbx * 16 + (hl & 1) * 8];
        const uint8_t *pl = (hl >> 4) ? d.cr : d.cb;
        f.chroma = *(const uint32_t *)&pl[(uint64_t)(mby * 8 + ((hl >> 1) & 7)) * d.wc + mbx * 8 + (hl & 1) * 4];
        if (hl < 8) f.bs = *(const uint32_t *)&G(P.bs)[(size_t)(mby * P.W + mbx) * 32 + hl * 4];
    }
    return f;
}
__device__ __forceinline__ void dbk_place(const DbkPrefetch &f, WaveDbk *wd, uint8_t *bsrec, int cur) {
    const int hl = half_lane();
    *(uint2 *)&wd->L[(hl >> 1) + 4][cur + (hl & 1) * 8] = f.luma;
    *(uint32_t *)&wd->C[hl >> 4][((hl >> 1) & 7) + 4][(cur >> 1) + (hl & 1) * 4] = f.chroma;
    if (hl < 8) *(uint32_t *)&bsrec[hl * 4] = f.bs;
}

// Rows 0 .. ry (chroma 0 .. cy) of the MB whose column 0 sits at tile column `col` to the
// picture: one aligned 16-byte luma row per half-lane 0..15, one 8-byte chroma row per
// half-lane 16..31.
__device__ __forceinline__ void dbk_store_mb(const DbkPlanes &d, int mbx, int mby, int col, int ry, int cy,
                                             const WaveDbk *wd) {
    const int hl = half_lane();
    if (hl < 16) {
        if (hl <= ry)
            *(uint4 *)&d.y[(uint64_t)(mby * 16 + hl) * d.wl + mbx * 16] = *(const uint4 *)&wd->L[hl + 4][col];
    } else {
        const int k = hl - 16, c = k >> 3, r = k & 7;
        if (r <= cy)
            *(uint2 *)&(c ? d.cr : d.cb)[(uint64_t)(mby * 8 + r) * d.wc + mbx * 8] = *(const uint2 *)&wd->C[c][r + 4][col >> 1];
    }
}
// Rows 13..15 (chroma 5..7) of the MB above the current one (tile column cur), final after
// the current MB's top edge: half-lanes 0..2 luma, 3..8 chroma.
__device__ __forceinline__ void dbk_store_top(const DbkPlanes &d, int mbx, int mby, const WaveDbk *wd, int cur) {
    const int hl = half_lane();
    if (hl < 3) {
        *(uint4 *)&d.y[(uint64_t)(mby * 16 - 3 + hl) * d.wl + mbx * 16] = *(const uint4 *)&wd->L[hl + 1][cur];
    } else if (hl < 9) {
        const int k = hl - 3, c = k >= 3, r = k - 3 * c;
        *(uint2 *)&(c ? d.cr : d.cb)[(uint64_t)(mby * 8 - 3 + r) * d.wc + mbx * 8] = *(const uint2 *)&wd->C[c][r + 1][cur >> 1];
    }
}

// A strip is one dword per half-lane: luma row 12 + j / 4, bytes 4 (j & 3) .. + 3 for
// j < 16; for j = 16 + k the chroma plane k / 8, row 4 + (k / 2) % 4, bytes 4 (k & 1) .. + 3.
// Strip dword of the MB at tile column `col` (rows 12..15 / chroma 4..7).
__device__ __forceinline__ uint32_t strip_word(const WaveDbk *wd, int col) {
    const int hl = half_lane();
    if (hl < 16) return *(const uint32_t *)&wd->L[16 + (hl >> 2)][col + (hl & 3) * 4];
    const int k = hl - 16, c = k >> 3, r = (k >> 1) & 3, h = k & 1;
    return *(const uint32_t *)&wd->C[c][8 + r][(col >> 1) + h * 4];
}
// A strip dword into the tile rows -4..-1 above the MB at tile column cur.
__device__ __forceinline__ void strip_place(WaveDbk *wd, int cur, uint32_t v) {
    const int hl = half_lane();
    if (hl < 16) {
        *(uint32_t *)&wd->L[hl >> 2][cur + (hl & 3) * 4] = v;
    } else {
        const int k = hl - 16, c = k >> 3, r = (k >> 1) & 3, h = k & 1;
        *(uint32_t *)&wd->C[c][r][(cur >> 1) + h * 4] = v;
    }
}

// Strips that cross from a band to the next (another workgroup, possibly another XCD) go
// through device memory as 32 words of 64 bits: the strip dword of half-lane j in the low
// half of word j, the launch epoch in the high half.  Each word is stored and loaded as
// one agent-scope atomic, so a reader that finds the epoch in all 32 words holds the
// whole strip; neither side needs a fence, whose L2 write-back / invalidation would also
// hit every other kernel running on those XCDs.
struct StripWords {
    uint64_t w[32];
};
__device__ __forceinline__ void strip_put_g(StripWords *sw, const WaveDbk *wd, int col, uint32_t epoch) {
    const uint32_t v = strip_word(wd, col);
    __hip_atomic_store(&sw->w[half_lane()], (uint64_t)epoch << 32 | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Waits (the whole wave) for the strip of this launch and puts it into the tile of the
// lanes where `mine` holds, above the MB at tile column cur.  Bounded like the other waits.
__device__ void strip_get_g(const StripWords *sw, WaveDbk *wd, int cur, uint32_t epoch, bool mine, uint32_t *err) {
    uint64_t v = (uint64_t)epoch << 32;
    for (uint32_t it = 0;; ++it) {
        if (mine) v = __hip_atomic_load(&sw->w[half_lane()], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!any_lane((uint32_t)(v >> 32) != epoch)) break;
        if (it >= SPIN_LIMIT) {
            atomicOr(err, DEVERR_TIMEOUT);
            break;
        }
        if ((it & 255u) == 255u && uniform(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0u) break;
        __builtin_amdgcn_s_sleep(1);
    }
    if (mine) strip_place(wd, cur, (uint32_t)v);
}

// Wave w of the band of rows y0 .. y1 - 1 (band index b): rows y0 + 2w and, when it lies
// in the band, y0 + 2w + 1.  Step s filters the upper row's MB s and the lower row's MB
// s - 1, then stores (and hands down) the MBs one to the left of those.
__device__ void deblock_pair_band(const DevPic &P, int w, int y0, int y1, int b, DbkWgShared &sh, uint32_t epoch,
                                  uint32_t *err) {
    const int W = P.W, H = P.H;
    const int yu = y0 + 2 * w, yl = yu + 1;
    const bool hasL = yl < y1, lower = wave_lane() >= 32;
    const int y = lower ? yl : yu;
    const DbkPlanes d = dbk_planes(P);
    WaveDbk *const tu = &sh.tiles[2 * w], *const tl = &sh.tiles[2 * w + 1];
    WaveDbk *const wd = lower ? tl : tu;
    uint8_t *const bsrec = sh.bsrec[2 * w + (lower ? 1 : 0)];
    const bool last = y == H - 1, topU = yu > 0, top = lower || topU;
    const int ry = last ? 15 : 12, cy = last ? 7 : 4;  // rows a half stores of its own MBs
    // the band's last row hands its strips to the next band (none after the picture's last)
    const bool gU = !hasL && yu == y1 - 1 && yu != H - 1, gL = hasL && yl == y1 - 1 && yl != H - 1;
    const bool ringL = hasL && yl + 1 < y1;  // the lower row hands down to wave w + 1
    const bool from_g = topU && w == 0;      // the upper row reads the band above's strips
    StripSlot *const ring = sh.ring[w];
    const StripSlot *const above = w > 0 ? sh.ring[w - 1] : nullptr;
    StripWords *const gout = (StripWords *)G(P.gstrip) + (size_t)b * W;              // this band -> next
    const StripWords *const gin = (const StripWords *)G(P.gstrip) + (size_t)(b - 1) * W;  // previous band -> this
    const int steps = W + (hasL ? 1 : 0);  // the lower row's last MB is stored in step W + 1
    DbkPrefetch f = dbk_fetch(P, d, 0, yu, !lower);
    // deblocking words of each half's MB and of the MB above it, one step ahead like the samples
    uint32_t qq = 0u, qn = lower ? 0u : G(P.dbq)[yu * W], tn = (!lower && topU) ? G(P.dbq)[(yu - 1) * W] : 0u;
    for (int s = 0; s <= steps; s++) {
        const int x = lower ? s - 1 : s;  // this half's MB
        const bool act = lower ? (hasL && s >= 1 && s <= W) : s < W;
        const int cur = 16 * (x & 1), prev = 16 - cur;  // tile columns of MB x and MB x - 1
        const int prevU = 16 - 16 * (s & 1);              // the upper row's MB s - 1 (= the lower's MB column)
        const uint32_t lq = qq, tq = tn;
        qq = qn;
        if (act) dbk_place(f, wd, bsrec, cur);
        {  // next step's MBs, in flight during this one
            const bool nU = s + 1 < W, nL = hasL && s < W;
            f = dbk_fetch(P, d, lower ? s : s + 1, y, lower ? nL : nU);
            const uint32_t qu = nU ? G(P.dbq)[yu * W + s + 1] : 0u, ql = nL ? G(P.dbq)[yl * W + s] : 0u;
            const uint32_t tu_ = nU && topU ? G(P.dbq)[(yu - 1) * W + s + 1] : 0u, tl_ = nL ? G(P.dbq)[yu * W + s] : 0u;
            qn = lower ? ql : qu;
            tn = lower ? tl_ : tu_;
        }
        wave_sync();
        dbk_lines<0, true>(P, lq, x > 0, qq, bsrec, wd, &sh.tb, cur, act);  // vertical edges
        wave_sync();
        // MB x - 1 is final now (its columns 13..15 took MB x's left edge) except rows 13..15,
        // which the next row's top edge filters: store it and hand its strip down before
        // waiting for the row above, so the row below runs about one step behind
        const bool pU = s >= 1 && s <= W, pL = hasL && s >= 2;
        if (lower ? pL : pU) dbk_store_mb(d, x - 1, y, prev, ry, cy, wd);
        if (hasL && pU && lower) strip_place(tl, prevU, strip_word(tu, prevU));  // upper -> lower
        if (gU && pU && !lower) strip_put_g(&gout[s - 1], tu, prevU, epoch);
        if (gL && pL && lower) strip_put_g(&gout[s - 2], tl, prev, epoch);
        if (ringL && pL) {
            const int m = s - 2;
            if (m - STRIP_SLOTS >= 0) lds_wait(&sh.cons[w], (uint32_t)(m - STRIP_SLOTS + 1), err);  // slot read?
            if (lower) {
                StripSlot *sl = &ring[m % STRIP_SLOTS];
                sl->w[half_lane()] = strip_word(tl, prev);
            }
            lds_post(&sh.prog[w], (uint32_t)m + 1);
        }
        if (topU && s < W) {  // the upper row's strip from the row above
            if (from_g) {
                strip_get_g(&gin[s], tu, 16 * (s & 1), epoch, !lower, err);
            } else {
                lds_wait(&sh.prog[w - 1], (uint32_t)s + 1, err);
                if (!lower) strip_place(tu, 16 * (s & 1), above[s % STRIP_SLOTS].w[half_lane()]);
                lds_post(&sh.cons[w - 1], (uint32_t)s + 1);
            }
        }
        wave_sync();
        dbk_lines<1, true>(P, tq, top, qq, bsrec, wd, &sh.tb, cur, act);  // horizontal edges
        wave_sync();
        if (act && top) dbk_store_top(d, x, y, wd, cur);
        // the next step's tile writes (the other half of the tile) follow this barrier
        wave_sync();
    }
}

// One workgroup per band of `rows` MB rows (rows y0 + 2w, y0 + 2w + 1 on wave w).
__global__ __launch_bounds__(WAVE * DBK_WG_MAX_WAVES) void k_deblock_band(const DevPic *__restrict__ pics,
                                                                          const uint32_t *__restrict__ items, uint32_t n,
                                                                          int rows, DevCtl *ctl, uint32_t *epoch_p,
                                                                          uint32_t *err) {
    __shared__ DbkWgShared sh;
    __shared__ uint32_t s_ticket;
    const uint32_t epoch = launch_epoch(epoch_p);
    const int nw = blockDim.x / WAVE, w = (int)uniform(threadIdx.x / WAVE);
    if ((int)threadIdx.x < nw) sh.cons[threadIdx.x] = 0;
    dbk_tables_init(&sh.tb);
    const uint32_t t = band_ticket(ctl, &s_ticket, sh.prog, nw);
    if (t < n) {
        const uint32_t it = items[t];
        const DevPic &P = pics[it >> WI_PIC_SHIFT];
        const int b = (int)(it & WI_MB_MASK), y0 = b * rows, y1 = min(P.H, y0 + rows);
        if (P.dbk_idc != 1 && y0 + 2 * w < y1) deblock_pair_band(P, w, y0, y1, b, sh, epoch, err);
    }
    band_retire(ctl, epoch_p, epoch);
}

// Boundary strengths of every edge segment of every MB (8.7.2.1, Deblock.cpp:994-1310),
// stored [direction][4-line segment][edge] so that a deblocking line reads its four edges
// as one dword, and each MB's deblocking word (G(P.dbq): qP of luma / Cb edges / Cr edges in
// bytes 0..2, transform_size_8x8_flag in bit 24, the own-qP flags of its MB edges in bits
// 25 / 26); fully parallel: they depend on syntax
// only, not on samples.  Half a wave per MB; lane j = 16 * dir + 4 * edge + segment.
__global__ __launch_bounds__(WAVE) void k_bs(const DevPic *__restrict__ pics, const uint32_t *__restrict__ items, uint32_t n) {
    const int lane = wave_lane();
    const uint32_t i = blockIdx.x * 2 + (lane >> 5);
    if (i >= n) return;
    const uint32_t it = items[i];
    const DevPic &P = pics[it >> WI_PIC_SHIFT];
    const int addr = (int)(it & WI_MB_MASK), W = P.W, mbx = addr % W, mby = addr / W;
    const int j = lane & 31, dir = j >> 4, e = ((j >> 2) & 3) * 4, k = (j & 3) * 4;
    int bS = 0;
    const bool has_nb = dir == 0 ? mbx > 0 : mby > 0;
    const uint32_t sq = P.multi_slice ? G(P.mbs)[addr].slice_num : 0u;
    const bool other = P.multi_slice && e == 0 && has_nb && G(P.mbs)[dir == 0 ? addr - 1 : addr - W].slice_num != sq;
    // An MB edge at the picture border is not filtered, nor, with disable_deblocking_filter_idc
    // 2, one between two slices (8.7: filterLeftMbEdgeFlag / filterTopMbEdgeFlag); no edge
    // of an MB at or after the deblocking stop is.  With idc 0 the reference filters an
    // edge between slices with the current MB in the neighbour's place (its neighbour
    // derivation reports the other slice's MB as unavailable and the p-MB address stays
    // the current one, Deblock.cpp:747-760, 784-791): the bS of an internal edge between
    // the current MB's blocks at x (or y) 15 and 0, and the current MB's qP on both sides.
    const bool own = e == 0 && other && P.dbk_idc == 0;
    const int ap = (e || own) ? addr : (dir == 0 ? addr - 1 : addr - W);
    const bool edge = (uint32_t)addr < P.dbk_stop && (e > 0 || (has_nb && !(P.dbk_idc == 2 && other)));
    if (edge) {
        MbView Qv, Pv;
        load_view(P, addr, Qv);
        load_view(P, ap, Pv);
        const int xq = dir == 0 ? e : k, yq = dir == 0 ? k : e;
        const int xp = dir == 0 ? (e ? e - 1 : 15) : k, yp = dir == 0 ? k : (e ? e - 1 : 15);
        bS = bs_luma(P, ap, Pv, addr, Qv, e == 0 && !own, xp, yp, xq, yq);
    }
    G(P.bs)[(size_t)addr * 32 + (j & 16) + ((j & 3) << 2) + ((j >> 2) & 3)] = (uint8_t)bS;  // [dir][segment][edge]
    if (j == 0) {
        const DevAux &ax = G(P.aux)[addr];
        const bool own_t = P.multi_slice && P.dbk_idc == 0 && mby > 0 && G(P.mbs)[addr - W].slice_num != sq;
        G(P.dbq)[addr] = (uint32_t)ax.qdb[0] | ((uint32_t)ax.qdb[1] << 8) | ((uint32_t)ax.qdb[2] << 16) |
                         ((ax.kind & AUX_T8) ? 1u << 24 : 0u) | (own ? DBQ_OWN_LEFT : 0u) | (own_t ? DBQ_OWN_TOP : 0u);
    }
}

// YUV420P -> BGR24 of one DPB slot (CH264PictureBase::convertYuv420pToBgr24,
// H264PictureBase.cpp:440-495): one thread per 2x2 luma block (one chroma sample);
// BT.601 integer coefficients over 1000 with C division (toward zero), clipped.
__global__ __launch_bounds__(256) void k_bgr24(const uint8_t *__restrict__ yuv, int wl, int hl, uint8_t *__restrict__ bgr,
                                                uint32_t stride, int flip) {
    const int wc = wl >> 1, hc = hl >> 1;
    const int cx = blockIdx.x * blockDim.x + threadIdx.x, cy = blockIdx.y;
    if (cx >= wc || cy >= hc) return;
    const uint8_t *u = yuv + (size_t)wl * hl, *v = u + (size_t)wc * hc;
    const int U = u[(size_t)cy * wc + cx] - 128, V = v[(size_t)cy * wc + cx] - 128;
    const int gb = 2018 * U, gg = -813 * V - 391 * U, gr = 1596 * V;
#pragma unroll
    for (int dy = 0; dy < 2; dy++) {
        const int y = 2 * cy + dy;
        uint8_t *o = bgr + (size_t)(flip ? hl - 1 - y : y) * stride + (size_t)(2 * cx) * 3;
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
            const int Y = 1164 * (yuv[(size_t)y * wl + 2 * cx + dx] - 16);
            o[3 * dx + 0] = (uint8_t)clip1((Y + gb) / 1000);
            o[3 * dx + 1] = (uint8_t)clip1((Y + gg) / 1000);
            o[3 * dx + 2] = (uint8_t)clip1((Y + gr) / 1000);
        }
    }
}

__global__ void k_copy(const uint4 *src, uint4 *dst, uint64_t n16) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

}  // namespace

extern "C" {

int h264g_launch_inter(const DevPic *pics, const uint32_t *items, uint32_t n, const uint32_t *intra_items, uint32_t n_intra,
                       int mbaff, int scaled, uint32_t *err, void *stream) {
    if (n + n_intra == 0) return 0;
    if (mbaff)  // (scaling matrices are refused in MBAFF frames)
        hipLaunchKernelGGL((k_inter<true, false>), dim3(n + n_intra), dim3(WAVE), 0, (hipStream_t)stream, pics, items, n,
                           intra_items, n_intra, err);
    else if (scaled)
        hipLaunchKernelGGL((k_inter<false, true>), dim3(n + n_intra), dim3(WAVE), 0, (hipStream_t)stream, pics, items, n,
                           intra_items, n_intra, err);
    else
        hipLaunchKernelGGL((k_inter<false, false>), dim3(n + n_intra), dim3(WAVE), 0, (hipStream_t)stream, pics, items, n,
                           intra_items, n_intra, err);
    return (int)hipGetLastError();
}

int h264g_launch_intra(const DevPic *pics, const uint32_t *items, uint32_t n, DevCtl *ctl, uint32_t *epoch, uint32_t *err,
                       int waves, int mbaff, void *stream) {
    if (n == 0) return 0;
    uint32_t g = (uint32_t)waves < n ? (uint32_t)waves : n;
    if (mbaff)
        hipLaunchKernelGGL(k_intra<true>, dim3(g), dim3(WAVE), 0, (hipStream_t)stream, pics, items, n, ctl, epoch, err);
    else
        hipLaunchKernelGGL(k_intra<false>, dim3(g), dim3(WAVE), 0, (hipStream_t)stream, pics, items, n, ctl, epoch, err);
    return (int)hipGetLastError();
}

int h264g_launch_deblock(const DevPic *pics, const uint32_t *items, uint32_t n, DevCtl *ctl, uint32_t *epoch, uint32_t *err,
                         int waves, void *stream) {
    if (n == 0) return 0;
    uint32_t g = (uint32_t)waves < n ? (uint32_t)waves : n;
    hipLaunchKernelGGL(k_deblock, dim3(g), dim3(WAVE), 0, (hipStream_t)stream, pics, items, n, ctl, epoch, err);
    return (int)hipGetLastError();
}

int h264g_launch_deblock_rows(const DevPic *pics, const uint32_t *items, uint32_t n, DevCtl *ctl, uint32_t *epoch,
                              uint32_t *err, int waves, void *stream) {
    if (n == 0) return 0;
    uint32_t g = (uint32_t)waves < n ? (uint32_t)waves : n;
    hipLaunchKernelGGL(k_deblock_rows, dim3(g), dim3(WAVE), 0, (hipStream_t)stream, pics, items, n, ctl, epoch, err);
    return (int)hipGetLastError();
}

int h264g_launch_intra_band(const DevPic *pics, const uint32_t *bands, uint32_t n, DevCtl *ctl, uint32_t *epoch,
                            int waves, uint32_t *err, void *stream) {
    if (n == 0) return 0;
    if (waves < 1 || waves > INTRA_WG_MAX_WAVES) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_intra_band, dim3(n), dim3(WAVE * waves), 0, (hipStream_t)stream, pics, bands, n, ctl, epoch, err);
    return (int)hipGetLastError();
}

int h264g_launch_deblock_band(const DevPic *pics, const uint32_t *bands, uint32_t n, DevCtl *ctl, uint32_t *epoch,
                              int rows, uint32_t *err, void *stream) {
    if (n == 0) return 0;
    if (rows < 1 || rows > 2 * DBK_WG_MAX_WAVES) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_deblock_band, dim3(n), dim3(WAVE * ((rows + 1) / 2)), 0, (hipStream_t)stream, pics, bands, n,
                       rows, ctl, epoch, err);
    return (int)hipGetLastError();
}

int h264g_launch_deblock_mbaff(const DevPic *pics, const uint32_t *items, uint32_t n, DevCtl *ctl, uint32_t *epoch,
                               uint32_t *err, int waves, void *stream) {
    if (n == 0) return 0;
    uint32_t g = (uint32_t)waves < n ? (uint32_t)waves : n;
    hipLaunchKernelGGL(k_deblock_mbaff, dim3(g), dim3(WAVE), 0, (hipStream_t)stream, pics, items, n, ctl, epoch, err);
    return (int)hipGetLastError();
}

int h264g_launch_bs(const DevPic *pics, const uint32_t *items, uint32_t n, void *stream) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_bs, dim3((n + 1) / 2), dim3(WAVE), 0, (hipStream_t)stream, pics, items, n);
    return (int)hipGetLastError();
}

int h264g_launch_bgr24(const uint8_t *yuv, int wl, int hl, uint8_t *bgr, uint32_t stride, int flip, void *stream) {
    const int wc = wl / 2, hc = hl / 2;
    hipLaunchKernelGGL(k_bgr24, dim3((wc + 255) / 256, hc), dim3(256), 0, (hipStream_t)stream, yuv, wl, hl, bgr, stride, flip);
    return (int)hipGetLastError();
}

int h264g_launch_copy(const uint8_t *src, uint8_t *dst, uint64_t bytes, void *stream) {
    uint64_t n16 = bytes / 16;
    if (n16 == 0) return 0;
    hipLaunchKernelGGL(k_copy, dim3(1024), dim3(256), 0, (hipStream_t)stream, (const uint4 *)src, (uint4 *)dst, n16);
    return (int)hipGetLastError();
}

}  // extern "C"

// ------------------------------------------------------------ static scratch
// Device memory of the allocation-free entry point (h264g_picture_decode_static): the
// staged syntax of one picture and its stage buffers (records, levels, work lists,
// boundary strengths, intra residuals).  Every device the library runs on holds its own
// copy until the process ends.  In host code `scratch` names a host-side stand-in, so its
// device address is looked up once per device and only that address is used.
namespace {
__device__ __align__(256) uint8_t scratch[H264G_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices] = {};
}  // namespace

int h264g_scratch_address(int device, void **out) {
    if (device < 0 || device >= kMaxDevices || !out) return (int)hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return (int)e;
        if (cur != device) return (int)hipErrorInvalidDevice;  // the caller makes `device` current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return (int)hipErrorInvalidSymbol;
        g_scratch_addr[device] = p;
    }
    *out = g_scratch_addr[device];
    return 0;
}
