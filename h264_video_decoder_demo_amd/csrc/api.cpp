// api.cpp -- host side of libh264g: the C ABI of include/h264g.h.
//
// Staging a picture (h264g_picture_begin) validates the host parser's syntax
// records, derives the per-MB values the kernels need (DPB slots of the
// reference indices, resolved weighted-prediction entries, non-zero masks for
// the boundary strength, chroma / deblocking QPs) and builds the three device
// work queues: inter MBs, intra MBs and all MBs, the latter two ordered by the
// wavefront time t = x + 2y so that every dependency of a queue entry precedes it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <vector>

#include "h264g.h"
#include "h264g_dev.h"
#include "h264g_internal.h"

namespace {

thread_local char g_err[512] = "";

int chroma_qp(int qpy, int offset) {  // 8.5.8, Table 8-15 (H264PictureBase.cpp:4748-4813)
    static const uint8_t tab[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};
    int qpi = std::min(51, std::max(0, qpy + offset));
    return qpi < 30 ? qpi : tab[qpi - 30];
}

bool cls_inter(int c) { return c == H264G_MB_P || c == H264G_MB_P_SKIP || c == H264G_MB_B; }
// An MB that constrained intra prediction hides from intra MBs.  The reference tests
// IS_INTER_Prediction_Mode(m_mb_pred_mode) (H264PictureBase.cpp:1129, 1492, 1905, 2158,
// 836, 973; H264CommonFunc.h:273), which holds for Pred_L0 / Pred_L1 / BiPred only: P_8x8
// and B_8x8 MBs have m_mb_pred_mode Pred_NA (H264MacroBlock.cpp:99-100, 142) and stay
// available to intra prediction; P_Skip is Pred_L0.
bool cip_hides(const h264g_mb &m) {
    return m.mb_type == H264G_MB_P_SKIP ||
           ((m.mb_type == H264G_MB_P || m.mb_type == H264G_MB_B) && m.part_shape != H264G_PART_8x8);
}
bool cls_intra(int c) { return c == H264G_MB_I4x4 || c == H264G_MB_I8x8 || c == H264G_MB_I16x16; }

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) return h264g_set_error(H264G_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct Stream {
    h264g_stream_desc desc;
    int W = 0, H = 0;
    uint8_t *dpb = nullptr;
    uint64_t slot_bytes = 0;
    // H264G_MAX_STREAM_PICS progress sets, one for each picture of the stream in one
    // launch, each: per-MB progress words | per-row progress words | deblocking strips
    // between row bands | intra strips | MBAFF streams: deblocking progress per MB pair
    uint8_t *prog = nullptr;
    uint64_t set_bytes = 0, o_rowprog = 0, o_gstrip = 0, o_istrip = 0, o_pair = 0;
    bool mbaff = false;            // MBAFF frames (pairs of field / frame MBs)
    int strip_dbk_rows = 0, strip_intra_rows = 0;  // strips sized for bands of at least these many rows
    uint64_t bytes = 0;
    int lane = 0;                  // HIP stream (lane) of the context that runs its pictures
    uint64_t gen = 0;              // open count of the context when this stream was opened:
                                   // pictures and batches staged for an earlier stream with the
                                   // same id carry another value and are refused
    bool open = false;
};

// Points a picture's progress words and strips at progress set k of its stream.
void use_set(const Stream &s, int k, DevPic &P) {
    uint8_t *b = s.prog + (uint64_t)k * s.set_bytes;
    P.flags = (uint32_t *)b;
    P.rowprog = (uint64_t *)(b + s.o_rowprog);
    P.gstrip = b + s.o_gstrip;
    P.istrip = (uint64_t *)(b + s.o_istrip);
    P.pairprog = s.mbaff ? (uint32_t *)(b + s.o_pair) : nullptr;
}

struct Batch;

// The DPB slots one picture writes (dst) and reads (its reference pictures).
struct SlotUse {
    uint32_t stream;
    int32_t dst;
    std::vector<int32_t> refs;
};
SlotUse slot_use(uint32_t stream, const h264g_pic_desc &d) {
    SlotUse u{stream, d.dst_slot, {}};
    for (int l = 0; l < 2; l++)
        for (int r = 0; r < 32; r++)
            if (d.ref_slot[l][r] >= 0 && std::find(u.refs.begin(), u.refs.end(), d.ref_slot[l][r]) == u.refs.end())
                u.refs.push_back(d.ref_slot[l][r]);
    return u;
}

// Deblocking launches of one lane still possibly in flight on its deblocking stream: the
// event recorded after each and the (stream, slot) pairs it writes.
constexpr int kDbkHist = 4;
struct DbkRec {
    hipEvent_t ev = nullptr;
    std::vector<std::pair<uint32_t, int32_t>> dst;
    bool live = false;
};

}  // namespace

struct h264g_pic {
    h264g_ctx *ctx = nullptr;
    uint32_t stream = 0;
    uint64_t stream_gen = 0;       // Stream::gen at staging: its DPB and strips are valid only
                                   // while the stream is open with that value
    h264g_pic_desc desc;
    int W = 0, H = 0;
    void *dmem = nullptr;          // one device block for everything below (pooled)
    size_t dbytes = 0;
    h264g_mb *d_mbs = nullptr;
    DevAux *d_aux = nullptr;
    int16_t *d_levels = nullptr;
    DevWeight *d_wt = nullptr;
    uint32_t *d_items = nullptr;   // inter | intra | deblock (per MB) | deblock (rows) queues
    DevPic *d_pic = nullptr;
    uint32_t n_inter = 0, n_intra = 0, n_mbs = 0, n_rows = 0;
    DevPic hpic;                   // host copy (batches re-pack it)
    std::vector<uint32_t> intra_t, all_t;  // host copies of the t-ordered queues (MB addresses;
                                           // MBAFF frames: intra MBs in pair t order, and
                                           // all_t holds MB pair indices)
    std::vector<uint32_t> inter_list;
    bool mbaff = false;            // an MBAFF frame: intra on the per-MB wavefront, deblocking
                                   // over MB pairs (the band schedules take frame MBs only)
    uint32_t n_pairs = 0;
    uint64_t alg[3] = {0, 0, 0};
};

namespace {
// The pictures of one batch on one lane.
struct Batch {
    void *dmem = nullptr;          // pooled device block
    size_t dbytes = 0;
    DevPic *d_pics = nullptr;
    uint32_t n_pics = 0;
    uint32_t *d_inter = nullptr, *d_intra = nullptr, *d_all = nullptr, *d_rows = nullptr;
    uint32_t *d_ibands = nullptr, *d_dbands = nullptr;   // (picture, band) items, band-major
    uint32_t n_inter = 0, n_intra = 0, n_all = 0, n_rows = 0, n_ibands = 0, n_dbands = 0;
    // MBAFF frames: inter MBs (after the frame pictures' in d_inter), intra MBs, MB pairs
    uint32_t *d_minter = nullptr, *d_mintra = nullptr, *d_mpairs = nullptr;
    uint32_t n_minter = 0, n_mintra = 0, n_mpairs = 0;
    bool dbk_off = false;            // no picture of the batch is deblocked
    bool scaled = false;             // a picture of the batch has scaling matrices
    hipGraphExec_t graph = nullptr;  // the captured stage sequence (kernel nodes only)
    hipGraphExec_t graph_dbk = nullptr;  // its deblocking, when that runs on the lane's deblocking stream
    std::vector<SlotUse> uses;       // DPB slots the pictures write and read
    std::vector<std::pair<uint32_t, uint64_t>> streams;  // (stream id, Stream::gen) of its pictures
};
}  // namespace

struct h264g_batch {
    Batch sub[2];                  // pictures of lane 0 / lane 1
    int dbk_rows = 0, intra_rows = 0;  // band heights its band lists were built for
};

struct h264g_ctx {
    int device = 0;
    // Two lanes (HIP streams): streams of pictures with at least split_mbs MBs run on lane
    // 1, the others on lane 0, so a batch's long intra / deblocking chains of large
    // pictures overlap the other pictures' stages instead of ending every stage.
    hipStream_t stream = nullptr;  // lane 0
    hipStream_t aux = nullptr;     // lane 1
    // With the band deblocking schedule, a lane's deblocking launches go to a stream of
    // their own (dstream), after the lane's other stages of the same pictures.  The next
    // pictures' stages wait for a deblocking launch only when they write or read a DPB
    // slot it writes (hist), so they overlap the deblocking of independent pictures (e.g.
    // a B picture not used for reference).
    hipStream_t dstream[2] = {nullptr, nullptr};
    hipEvent_t front[2] = {nullptr, nullptr};  // end of a lane's last pre-deblocking stages
    DbkRec hist[2][kDbkHist];
    int hist_next[2] = {0, 0};
    int split_mbs = 16384;         // 0: one lane
    uint32_t *d_err = nullptr;
    DevCtl *d_ctl = nullptr;       // per lane L: [4L] intra, [4L + 1] deblock, [4L + 2].ticket = intra launch
                                   // epoch, [4L + 3].ticket = deblocking launch epoch
    int waves = 2048;
    uint64_t budget = 0, used = 0;
    std::vector<Stream> streams;
    uint64_t opened = 0;           // streams opened so far (Stream::gen)
    // Device blocks of released pictures / batches, reusable once `ready` has completed on
    // the stream (so releasing never waits for the GPU and staging rarely allocates).
    struct PoolBlock { void *ptr; size_t size; hipEvent_t ready[4]; };  // lanes, then their deblocking streams
    std::vector<PoolBlock> pool;
    size_t pool_bytes = 0;
    bool profile = false;
    bool no_graphs = false;
    int dbk_schedule = H264G_DEBLOCK_BANDS;
    int intra_schedule = H264G_INTRA_BANDS;
    int intra_waves = 8;           // rows per band (waves per workgroup) of the intra band schedule
    int dbk_waves = 8;             // MB rows per band of the deblocking band schedule (two per wave)
    std::vector<hipEvent_t> ev;    // 6 events per profiled run: inter | intra | bS | end (lane), deblock | end
    size_t ev_used = 0;            // events recorded since the last h264g_stage_times
    std::mutex mu;
};

extern "C" int32_t h264g_set_error(int32_t code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

namespace {

hipStream_t lane_stream(const h264g_ctx *c, int lane) { return lane ? c->aux : c->stream; }
// The four streams: lane 0, lane 1, their deblocking streams.
hipStream_t any_stream(const h264g_ctx *c, int i) { return i < 2 ? lane_stream(c, i) : c->dstream[i - 2]; }

hipError_t sync_lanes(h264g_ctx *c) {
    hipError_t r = hipSuccess;
    for (int i = 0; i < 4; i++) {
        const hipError_t e = hipStreamSynchronize(any_stream(c, i));
        if (r == hipSuccess) r = e;
    }
    return r;
}
// A lane and its deblocking stream.
hipError_t sync_lane(h264g_ctx *c, int lane) {
    const hipError_t e = hipStreamSynchronize(lane_stream(c, lane));
    const hipError_t e2 = hipStreamSynchronize(c->dstream[lane]);
    return e != hipSuccess ? e : e2;
}

// A picture staged for stream `id` opened as `gen` may still be launched: the stream is
// open and is the same one (a closed stream's DPB, progress words and strips are back in
// the pool, and its id may belong to a newer stream).  Caller holds ctx->mu.
int check_stream(const h264g_ctx *c, uint32_t id, uint64_t gen) {
    if (id >= c->streams.size() || !c->streams[id].open || c->streams[id].gen != gen)
        return h264g_set_error(H264G_E_STATE, "stream %u was closed after the picture was staged", id);
    return 0;
}

// ------------------------------------------------------------ block pool
constexpr size_t kPoolMaxBytes = 8ull << 30;

size_t pool_class(size_t n) {  // sizes are rounded up so that blocks can be reused
    size_t c = 64 << 10;
    while (c < n && c < (64ull << 20)) c <<= 1;
    return c >= n ? c : (n + (16ull << 20) - 1) & ~((16ull << 20) - 1);
}

// A device block of at least n bytes (caller holds ctx->mu).
void *pool_get(h264g_ctx *c, size_t n) {
    const size_t want = pool_class(n);
    for (size_t i = 0; i < c->pool.size(); i++) {
        auto &b = c->pool[i];
        bool ready = b.size == want;
        for (int k = 0; k < 4 && ready; k++) ready = hipEventQuery(b.ready[k]) == hipSuccess;
        if (ready) {
            void *p = b.ptr;
            for (auto ev : b.ready) (void)hipEventDestroy(ev);
            c->pool_bytes -= b.size;
            c->pool.erase(c->pool.begin() + (long)i);
            return p;
        }
    }
    void *p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) return nullptr;
    return p;
}

// Returns a block; it becomes reusable when the work enqueued so far on all four streams
// has completed.
void pool_put(h264g_ctx *c, void *p, size_t n) {
    if (!p) return;
    hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ok = true;
    for (int l = 0; l < 4; l++)
        ok = ok && hipEventCreateWithFlags(&e[l], hipEventDisableTiming) == hipSuccess &&
             hipEventRecord(e[l], any_stream(c, l)) == hipSuccess;
    if (!ok) {
        (void)sync_lanes(c);  // cannot track it: free it once idle
        for (auto ev : e) if (ev) (void)hipEventDestroy(ev);
        (void)hipFree(p);
        return;
    }
    c->pool.push_back({p, pool_class(n), {e[0], e[1], e[2], e[3]}});
    c->pool_bytes += pool_class(n);
    while (c->pool_bytes > kPoolMaxBytes && !c->pool.empty()) {  // trim the oldest
        auto &b = c->pool.front();
        for (auto ev : b.ready) {
            (void)hipEventSynchronize(ev);
            (void)hipEventDestroy(ev);
        }
        (void)hipFree(b.ptr);
        c->pool_bytes -= b.size;
        c->pool.erase(c->pool.begin());
    }
}

// --------------------------------------------------------------- weights
// 8.4.2.3 mode selection and 8.4.3 values (H264InterPrediction.cpp:2526-2613, 2833-3050).
// Explicit entries read the tables at refIdx -1 for an unused list, which in the
// reference aliases the preceding slice-header member: luma_weight_lX[-1] is
// luma_weight_lX_flag (the last one parsed) and luma_offset_lX[-1] is
// luma_weight_lX[31] (H264SliceHeader.h:78-80).  L1-only luma uses that (w0, o0).
int derive_weight(const h264g_pic_desc &d, int r0, int r1, DevWeight &w) {
    std::memset(&w, 0, sizeof(w));
    int pf0 = r0 >= 0, pf1 = r1 >= 0;
    if (d.slice_type == H264G_SLICE_P) w.mode = d.weighted_pred_flag == 1;
    else if (d.slice_type == H264G_SLICE_B) w.mode = d.weighted_bipred_idc == 1 || (d.weighted_bipred_idc == 2 && pf0 && pf1);
    if (!w.mode) return 0;
    if (d.slice_type == H264G_SLICE_B && d.weighted_bipred_idc == 2) {
        int poc0 = d.ref_poc[0][r0], poc1 = d.ref_poc[1][r1];
        int ww0 = 32, ww1 = 32;
        // POC1 == POC0: the reference divides by zero (H264InterPrediction.cpp:2958);
        // the spec's 32/32 fallback applies.
        if (poc1 != poc0 && !d.ref_long_term[0][r0] && !d.ref_long_term[1][r1]) {
            int tb = std::min(127, std::max(-128, d.poc - poc0)), td = std::min(127, std::max(-128, poc1 - poc0));
            int tx = (16384 + std::abs(td / 2)) / td;
            int dsf = std::min(1023, std::max(-1024, (tb * tx + 32) >> 6));
            if (!((dsf >> 2) < -64 || (dsf >> 2) > 128)) { ww0 = 64 - (dsf >> 2); ww1 = dsf >> 2; }
        }
        for (int c = 0; c < 3; c++) { w.logWD[c] = 5; w.w0[c] = (int16_t)ww0; w.w1[c] = (int16_t)ww1; }
        return 0;
    }
    w.logWD[0] = (int16_t)d.luma_log2_weight_denom;
    w.logWD[1] = w.logWD[2] = (int16_t)d.chroma_log2_weight_denom;
    w.w0[0] = (int16_t)(pf0 ? d.luma_weight[0][r0] : d.luma_weight_flag_last[0]);
    w.o0[0] = (int16_t)(pf0 ? d.luma_offset[0][r0] : d.luma_weight[0][31]);
    w.w1[0] = (int16_t)(pf1 ? d.luma_weight[1][r1] : d.luma_weight_flag_last[1]);
    w.o1[0] = (int16_t)(pf1 ? d.luma_offset[1][r1] : d.luma_weight[1][31]);
    for (int c = 0; c < 2; c++) {
        w.w0[1 + c] = pf0 ? d.chroma_weight[0][r0][c] : 0;
        w.o0[1 + c] = pf0 ? d.chroma_offset[0][r0][c] : 0;
        w.w1[1 + c] = pf1 ? d.chroma_weight[1][r1][c] : 0;
        w.o1[1 + c] = pf1 ? d.chroma_offset[1][r1][c] : 0;
    }
    if (pf0 && pf1) {  // H264InterPrediction.cpp:3038-3047 (its Cr test mixes in w0Cb)
        auto bad = [](int s, int lw) { return !(-128 <= s && s <= (lw == 7 ? 127 : 128)); };
        if (bad(w.w0[0] + w.w1[0], w.logWD[0]) || bad(w.w0[1] + w.w1[1], w.logWD[1]) ||
            !(-128 <= w.w0[2] + w.w1[2] && w.w0[1] + w.w1[2] <= (w.logWD[2] == 7 ? 127 : 128)))
            return h264g_set_error(H264G_E_SYNTAX, "explicit bi-prediction weight sum out of range");
    }
    return 0;
}

// Mask index -> level slot, the order residual() parses the blocks (H264MacroBlock.cpp:1592-1856).
int level_slot(int i, int i16, int cbp) {
    int idx = 0;
    if (i16) { if (i == idx) return 0; idx++; }
    for (int q = 0; q < 4; q++)
        if (cbp & (1 << q))
            for (int b = 0; b < 4; b++) { if (i == idx) return 1 + 4 * q + b; idx++; }
    int cc = cbp >> 4;
    if (cc >= 1) for (int c = 0; c < 2; c++) { if (i == idx) return 17 + c; idx++; }
    if (cc == 2) for (int k = 0; k < 8; k++) { if (i == idx) return 19 + k; idx++; }
    return -1;
}
int count_masks(int i16, int cbp) {
    int n = i16 ? 1 : 0;
    for (int q = 0; q < 4; q++) if (cbp & (1 << q)) n += 4;
    if ((cbp >> 4) >= 1) n += 2;
    if ((cbp >> 4) == 2) n += 8;
    return n;
}

// Intra prediction modes must only use available neighbours; the reference leaves
// the block unwritten otherwise (H264PictureBase.cpp:1176-1186), which has no defined
// output, so such syntax is rejected.
bool nxn_mode_ok(int mode, bool left, bool top, bool tl) {
    switch (mode) {
    case 0: case 3: case 7: return top;
    case 1: case 8: return left;
    case 2: return true;
    case 4: case 5: case 6: return top && left && tl;
    default: return false;
    }
}

// Partitions of an inter MB as (x, y, w, h) in 4x4-block units (mb_pred /
// sub_mb_pred shapes, Tables 7-13 / 7-17).  Returns their number.
int partitions(const h264g_mb &m, int out[16][4]) {
    auto put = [&](int &n, int x, int y, int w, int h) { out[n][0] = x; out[n][1] = y; out[n][2] = w; out[n][3] = h; n++; };
    int n = 0;
    const int shape = m.mb_type == H264G_MB_P_SKIP ? H264G_PART_16x16 : m.part_shape;
    switch (shape) {
    case H264G_PART_16x16: put(n, 0, 0, 4, 4); break;
    case H264G_PART_16x8: put(n, 0, 0, 4, 2); put(n, 0, 2, 4, 2); break;
    case H264G_PART_8x16: put(n, 0, 0, 2, 4); put(n, 2, 0, 2, 4); break;
    default:
        for (int q = 0; q < 4; q++) {
            const int x0 = (q & 1) * 2, y0 = (q >> 1) * 2;
            switch (m.sub_part_shape[q]) {
            case H264G_SUB_8x8: put(n, x0, y0, 2, 2); break;
            case H264G_SUB_8x4: put(n, x0, y0, 2, 1); put(n, x0, y0 + 1, 2, 1); break;
            case H264G_SUB_4x8: put(n, x0, y0, 1, 2); put(n, x0 + 1, y0, 1, 2); break;
            default: for (int k = 0; k < 4; k++) put(n, x0 + (k & 1), y0 + (k >> 1), 1, 1); break;
            }
        }
    }
    return n;
}

// Validates and stages one picture.  Its device block comes from the context's pool, or,
// when `fixed` is given, is the caller's block of `fixed_cap` bytes (pictures that need
// more are refused).
int stage_picture(h264g_ctx *ctx, Stream &st, const h264g_pic_desc &d, const h264g_mb *mbs, uint32_t n_mbs,
                  const int16_t *levels, uint64_t n_levels, h264g_pic *pic, void *fixed = nullptr,
                  size_t fixed_cap = 0) {
    const int W = st.W, H = st.H;
    // Field pictures (PAFF): the reference decoder dereferences NULL building the
    // reference list of every P / B field (H264RefPicList.cpp:1197-1230) and misplaces
    // the MBs of I fields (field view at twice the pitch, H264PictureBase.cpp:205-208),
    // so there is no reference output for them to equal.
    if (d.structure != H264G_FRAME)
        return h264g_set_error(H264G_E_ARG, "field pictures (structure %d) are not supported", d.structure);
    if ((d.mbaff != 0) != st.mbaff)
        return h264g_set_error(H264G_E_ARG, "picture mbaff %d on a stream opened with mbaff %d", d.mbaff, (int)st.mbaff);
    const bool mbaff = st.mbaff;
    if (mbaff && d.constrained_intra_pred_flag)
        return h264g_set_error(H264G_E_ARG, "constrained intra prediction in MBAFF frames is not supported");
    if (d.slice_type < 0 || d.slice_type > 2) return h264g_set_error(H264G_E_ARG, "slice_type %d", d.slice_type);
    if ((int)n_mbs != W * H) return h264g_set_error(H264G_E_ARG, "n_mbs %u != %d", n_mbs, W * H);
    const int nslots = st.desc.dpb_slots;
    if (d.dst_slot < 0 || d.dst_slot >= nslots) return h264g_set_error(H264G_E_ARG, "dst_slot %d", d.dst_slot);
    for (int l = 0; l < 2; l++) {
        if (d.num_ref_idx_active[l] < 0 || d.num_ref_idx_active[l] > 32) return h264g_set_error(H264G_E_ARG, "num_ref_idx_active");
        for (int r = 0; r < d.num_ref_idx_active[l]; r++)
            if (d.ref_slot[l][r] < 0 || d.ref_slot[l][r] >= nslots || d.ref_slot[l][r] == d.dst_slot)
                return h264g_set_error(H264G_E_ARG, "ref_slot[%d][%d] = %d", l, r, d.ref_slot[l][r]);
    }
    if (d.disable_deblocking_filter_idc < 0 || d.disable_deblocking_filter_idc > 2 || d.filter_offset_a < -12 || d.filter_offset_a > 12 ||
        d.filter_offset_b < -12 || d.filter_offset_b > 12)
        return h264g_set_error(H264G_E_ARG, "deblocking parameters");
    // Slices.  Deblocking stops at the first MB of no slice (H264G_SLICE_NONE), where the
    // reference's bS derivation fails on the MB's reset type and the failure is ignored
    // (H264InterPrediction.cpp:3182-3186, H264PictureBase.cpp:707-708).  MBAFF frames take
    // one slice.
    uint32_t dbk_stop = n_mbs, first_sn = H264G_SLICE_NONE;
    bool multi_slice = false;
    for (uint32_t a = 0; a < n_mbs; a++) {
        const uint32_t sn = mbs[a].slice_num;
        if (sn == H264G_SLICE_NONE) {
            if (mbaff) return h264g_set_error(H264G_E_ARG, "mb %u: MBs of no slice in an MBAFF frame", a);
            if (dbk_stop == n_mbs) dbk_stop = a;
            continue;
        }
        if (first_sn == H264G_SLICE_NONE) first_sn = sn;
        if (sn != first_sn && mbaff) return h264g_set_error(H264G_E_ARG, "mb %u: several slices in an MBAFF frame", a);
        multi_slice |= sn != first_sn;
    }
    // Scaling matrices: LevelScale4x4 / 8x8 (8-315 / 8-318) of the two luma lists the
    // reference dequantises with (chroma takes the MB's luma list, quirk Q13).
    std::vector<uint16_t> lsv;
    if (d.scaling_matrix) {
        // the reference maps a field MB's lists with the field scan (H264PictureBase.cpp:4896-4897)
        if (mbaff) return h264g_set_error(H264G_E_ARG, "scaling matrices in MBAFF frames are not supported");
        static const int v4[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16}, {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
        static const int v8[6][6] = {{20, 18, 32, 19, 25, 24}, {22, 19, 35, 21, 28, 26}, {26, 23, 42, 24, 33, 31},
                                     {28, 25, 45, 26, 35, 33}, {32, 28, 51, 30, 40, 38}, {36, 32, 58, 34, 46, 43}};
        lsv.assign(LS_ENTRIES, 0);
        for (int l = 0; l < 2; l++)
            for (int m = 0; m < 6; m++) {
                for (int k = 0; k < 16; k++) {
                    const int i = k >> 2, j = k & 3, c = ((i | j) & 1) == 0 ? 0 : ((i & j) & 1) ? 1 : 2;
                    lsv[l * LS4_LIST + m * 16 + k] = (uint16_t)(d.weight_scale4x4[3 * l][k] * v4[m][c]);
                }
                for (int k = 0; k < 64; k++) {
                    const int i = k >> 3, j = k & 7;
                    const int c = (i % 4 == 0 && j % 4 == 0) ? 0 : (i % 2 == 1 && j % 2 == 1) ? 1
                                : (i % 4 == 2 && j % 4 == 2) ? 2 : ((i % 4 == 0 && j % 2 == 1) || (i % 2 == 1 && j % 4 == 0)) ? 3
                                : ((i % 4 == 0 && j % 4 == 2) || (i % 4 == 2 && j % 4 == 0)) ? 4 : 5;
                    lsv[LS8_AT + l * LS8_LIST + m * 64 + k] = (uint16_t)(d.weight_scale8x8[l][k] * v8[m][c]);
                }
            }
    }
    std::vector<DevAux> aux(n_mbs);
    std::vector<int8_t> bsref(mbaff ? (size_t)n_mbs * 8 : 0, (int8_t)-1);
    std::vector<DevWeight> wt;
    std::map<int, int> wmap;
    {
        DevWeight w0;
        std::memset(&w0, 0, sizeof(w0));
        wt.push_back(w0);  // entry 0: intra / unused
    }
    uint64_t alg_inter = 0, alg_intra = 0, n_part_lists = 0;
    for (uint32_t a = 0; a < n_mbs; a++) {
        const h264g_mb &m = mbs[a];
        DevAux &x = aux[a];
        std::memset(&x, 0, sizeof(x));
        for (int l = 0; l < 2; l++) for (int q = 0; q < 4; q++) x.refid[l][q] = -1;
        const int mbx = a % W, mby = a / W, cls = m.mb_type;
        const bool field = mbaff && (m.flags & H264G_MBF_FIELD);
        if (cls > H264G_MB_B) return h264g_set_error(H264G_E_ARG, "mb %u: mb_type %d", a, cls);
        if (m.qp_y > 51) return h264g_set_error(H264G_E_ARG, "mb %u: qp %d", a, m.qp_y);
        if (m.level_offset > n_levels) return h264g_set_error(H264G_E_ARG, "mb %u: level_offset", a);
        const int cbp = m.cbp, i16 = cls == H264G_MB_I16x16;
        const int t8 = (m.flags & H264G_MBF_T8x8) ? 1 : 0;
        if ((cbp & 15) && (cbp >> 4) > 2) return h264g_set_error(H264G_E_ARG, "mb %u: cbp", a);
        if ((cbp >> 4) > 2) return h264g_set_error(H264G_E_ARG, "mb %u: cbp", a);
        if (i16 && (cbp & 15) != 0 && (cbp & 15) != 15) return h264g_set_error(H264G_E_ARG, "mb %u: Intra16x16 cbp", a);
        if (cls == H264G_MB_P_SKIP && cbp) return h264g_set_error(H264G_E_ARG, "mb %u: P_Skip with cbp", a);
        if (cls == H264G_MB_I8x8 && !t8) return h264g_set_error(H264G_E_ARG, "mb %u: I8x8 without transform_size_8x8_flag", a);
        if ((cls == H264G_MB_I4x4 || cls == H264G_MB_I16x16) && t8)
            return h264g_set_error(H264G_E_ARG, "mb %u: transform_size_8x8_flag on an MB type that has no 8x8 transform", a);
        // transform_size_8x8_flag is only parsed for inter MBs with luma residual and
        // no sub-8x8 partitions (7.3.5, H264MacroBlock.cpp:1069-1071).
        if (cls_inter(cls) && t8 && ((cbp & 15) == 0 || (m.part_shape == H264G_PART_8x8 &&
                                                        (m.sub_part_shape[0] | m.sub_part_shape[1] | m.sub_part_shape[2] | m.sub_part_shape[3]))))
            return h264g_set_error(H264G_E_ARG, "mb %u: transform_size_8x8_flag without luma residual or with sub-8x8 partitions", a);
        if (!mbaff && (m.flags & H264G_MBF_FIELD)) return h264g_set_error(H264G_E_ARG, "mb %u: field MB in a frame picture", a);
        if (mbaff && ((m.flags ^ mbs[a ^ 1u].flags) & H264G_MBF_FIELD))
            return h264g_set_error(H264G_E_ARG, "mb %u: the MBs of a pair differ in mb_field_decoding_flag", a);
        if (d.slice_type == H264G_SLICE_I && cls_inter(cls)) return h264g_set_error(H264G_E_ARG, "mb %u: inter MB in I slice", a);
        if (d.slice_type != H264G_SLICE_B && cls == H264G_MB_B) return h264g_set_error(H264G_E_ARG, "mb %u: B MB outside B slice", a);
        if (d.slice_type != H264G_SLICE_P && (cls == H264G_MB_P || cls == H264G_MB_P_SKIP)) return h264g_set_error(H264G_E_ARG, "mb %u: P MB outside P slice", a);
        x.qdb[0] = (uint8_t)(cls == H264G_MB_IPCM ? 0 : m.qp_y);
        // Deblocking swaps the two chroma offsets (Deblock.cpp:813-815 passes iCbCr where
        // isChromaCb is expected; H264PictureBase.cpp:4816-4829).
        x.qdb[1] = (uint8_t)chroma_qp(x.qdb[0], d.second_chroma_qp_index_offset);
        x.qdb[2] = (uint8_t)chroma_qp(x.qdb[0], d.chroma_qp_index_offset);
        x.qpc[0] = (uint8_t)chroma_qp(m.qp_y, d.chroma_qp_index_offset);
        x.qpc[1] = (uint8_t)chroma_qp(m.qp_y, d.second_chroma_qp_index_offset);
        if (cls == H264G_MB_IPCM) {
            if (m.level_offset + 384 > n_levels) return h264g_set_error(H264G_E_ARG, "mb %u: PCM samples", a);
            x.kind = AUX_PCM;
            // RefIdxLX[0] = 0 and PredFlagLX = 0 from the zeroed MB state (Deblock.cpp:1175-1183)
            for (int l = 0; l < 2; l++)
                for (int q = 0; q < 4; q++) x.refid[l][q] = (int8_t)(d.num_ref_idx_active[l] > 0 ? d.ref_slot[l][0] : -1);
            if (mbaff) std::memcpy(&bsref[(size_t)a * 8], x.refid, 8);
            alg_inter += 384 + 384 * 2 + 16;
            continue;
        }
        if (cls_intra(cls)) x.kind |= AUX_INTRA;
        else x.kind |= AUX_INTER;
        if (t8) x.kind |= AUX_T8;
        // ---- levels: masks -> non-zero bits
        int nmask = count_masks(i16, cbp);
        if (cls == H264G_MB_P_SKIP) nmask = 0;
        x.nmask = (uint8_t)nmask;
        uint64_t nvals = 0;
        if (m.level_offset + (uint64_t)nmask > n_levels) return h264g_set_error(H264G_E_ARG, "mb %u: masks beyond levels", a);
        for (int i = 0; i < nmask; i++) {
            uint16_t mk = (uint16_t)levels[m.level_offset + i];
            int slot = level_slot(i, i16, cbp);
            int bits = (slot == 17 || slot == 18) ? 0x000F : ((slot == 0 || (slot >= 1 && slot <= 16 && !i16)) ? 0xFFFF : 0xFFFE);
            if (slot >= 1 && slot <= 16 && x.kind & AUX_T8) bits = 0xFFFF;
            if (mk & ~bits) return h264g_set_error(H264G_E_ARG, "mb %u: level mask %d has positions the block lacks", a, i);
            nvals += __builtin_popcount(mk);
            if (slot >= 1 && slot <= 16 && mk) {
                if (x.kind & AUX_T8) x.nz8 |= (uint8_t)(1 << ((slot - 1) >> 2));
                else x.nz4 |= (uint16_t)(1 << (slot - 1));
            }
        }
        if (m.level_offset + nmask + nvals > n_levels) return h264g_set_error(H264G_E_ARG, "mb %u: values beyond levels", a);
        x.ires = (uint32_t)nvals << AUX_NVALS_SHIFT;  // at most 27 x 16
        if (nmask) x.kind |= AUX_RESID;
        // ---- intra modes vs neighbour availability
        if (cls_intra(cls)) {
            auto avail = [&](int dx, int dy) {
                int nx = mbx + dx, ny = mby + dy;
                if (nx < 0 || ny < 0 || nx >= W) return false;
                if (mbs[ny * W + nx].slice_num != m.slice_num) return false;  // another slice (6.4.8)
                if (d.constrained_intra_pred_flag && cip_hides(mbs[ny * W + nx])) return false;
                return true;
            };
            bool A = avail(-1, 0), B = avail(0, -1), D = avail(-1, -1);
            if (mbaff) {  // 6.4.10 / Table 6-4: a frame MB at the bottom of its pair has its top
                          // MB above and the left pair above left; other MBs the pairs around
                const int pair = (int)(a >> 1), px = pair % W, py = pair / W;
                const bool in_pair = !field && (a & 1);
                A = px > 0;
                B = in_pair || py > 0;
                D = in_pair ? px > 0 : (py > 0 && px > 0);
            }
            if (cls == H264G_MB_I4x4 || cls == H264G_MB_I8x8) {
                int n = cls == H264G_MB_I4x4 ? 4 : 2;
                for (int by = 0; by < n; by++)
                    for (int bx = 0; bx < n; bx++) {
                        bool left = bx > 0 || A, top = by > 0 || B;
                        bool tl = (bx > 0 && by > 0) || (bx == 0 && by > 0 && A) || (bx > 0 && by == 0 && B) || (bx == 0 && by == 0 && D);
                        int blk = cls == H264G_MB_I4x4 ? (8 * (by / 2) + 4 * (bx / 2) + 2 * (by % 2) + (bx % 2)) : by * 2 + bx;
                        if (!nxn_mode_ok(m.intra_pred_modes[blk], left, top, tl))
                            return h264g_set_error(H264G_E_SYNTAX, "mb %u: intra mode %d of block %d uses unavailable samples", a, m.intra_pred_modes[blk], blk);
                    }
            } else {
                int md = m.intra16x16_pred_mode;
                if (md > 3 || (md == 0 && !B) || (md == 1 && !A) || (md == 3 && !(A && B && D)))
                    return h264g_set_error(H264G_E_SYNTAX, "mb %u: Intra16x16 mode %d", a, md);
            }
            int cm = m.intra_chroma_pred_mode;
            if (cm > 3 || (cm == 1 && !A) || (cm == 2 && !B) || (cm == 3 && !(A && B && D)))
                return h264g_set_error(H264G_E_SYNTAX, "mb %u: intra chroma mode %d", a, cm);
            alg_intra += 384 + 2 * (uint64_t)(nmask + nvals) + 16 * 2 + 48;
            continue;
        }
        // ---- inter: reference slots, prediction flags, weight entries
        if (m.part_shape > 3) return h264g_set_error(H264G_E_ARG, "mb %u: part_shape", a);
        if (m.part_shape == H264G_PART_8x8)
            for (int q = 0; q < 4; q++)
                if (m.sub_part_shape[q] > 3) return h264g_set_error(H264G_E_ARG, "mb %u: sub_part_shape", a);
        for (int q = 0; q < 4; q++) {
            int r[2] = {m.ref_idx[0][q], m.ref_idx[1][q]};
            if (r[0] < 0 && r[1] < 0) return h264g_set_error(H264G_E_ARG, "mb %u: quadrant %d predicts from no list", a, q);
            if (cls != H264G_MB_B && r[1] >= 0) return h264g_set_error(H264G_E_ARG, "mb %u: list 1 in a P MB", a);
            // field MBs of MBAFF frames: refIdx addresses the field list, the fields of
            // RefPicListX[refIdx / 2] (same parity as the MB when even), and weights sit at
            // refIdx >> 1 (H264InterPrediction.cpp:2169-2197, 2993-3031)
            const int nfield = field ? 2 : 1;
            for (int l = 0; l < 2; l++) {
                if (r[l] < -1 || r[l] >= d.num_ref_idx_active[l] * nfield)
                    return h264g_set_error(H264G_E_ARG, "mb %u: ref_idx[%d][%d] = %d", a, l, q, r[l]);
                if (r[l] >= 0) {
                    x.refid[l][q] = (int8_t)d.ref_slot[l][field ? r[l] >> 1 : r[l]];
                    x.pf |= (uint8_t)(1 << (4 * l + q));
                    if (field && ((r[l] & 1) ^ (int)(a & 1))) x.fpar |= (uint8_t)(1 << (4 * l + q));
                    // bS compares RefPicListX[RefIdxLX] even for field MBs (quirk Q4)
                    if (mbaff) bsref[(size_t)a * 8 + 4 * l + q] = (int8_t)d.ref_slot[l][r[l]];
                }
            }
            if (field && d.slice_type == H264G_SLICE_B && d.weighted_bipred_idc == 2 && r[0] >= 0 && r[1] >= 0)
                return h264g_set_error(H264G_E_ARG, "mb %u: implicit weights of field MBs are not supported", a);
            const int wr0 = field && r[0] >= 0 ? r[0] >> 1 : r[0], wr1 = field && r[1] >= 0 ? r[1] >> 1 : r[1];
            int key = (wr0 + 1) * 64 + (wr1 + 1);
            auto it = wmap.find(key);
            if (it == wmap.end()) {
                DevWeight w;
                int rc = derive_weight(d, wr0, wr1, w);
                if (rc) return rc;
                it = wmap.emplace(key, (int)wt.size()).first;
                wt.push_back(w);
                if (wt.size() > 255) return h264g_set_error(H264G_E_ARG, "too many weight entries");
            }
            x.wp[q] = (uint8_t)it->second;
        }
        // ---- algorithmic bytes of this MB: per partition and list used, the luma
        // footprint (w + 5) x (h + 5) with the 6-tap halo where the MV is fractional,
        // the two bilinear chroma footprints, the MB's output and its levels, and the
        // compact per-partition motion record (8 + 5 bytes per partition-list).
        int parts[16][4];
        const int np = partitions(m, parts);
        for (int k = 0; k < np; k++) {
            const int px = parts[k][0], py = parts[k][1], pw = parts[k][2], ph = parts[k][3];
            const int b = py * 4 + px, q = (py >> 1) * 2 + (px >> 1);
            for (int l = 0; l < 2; l++) {
                if (m.ref_idx[l][q] < 0) continue;
                const int w = 4 * pw, h = 4 * ph;
                const int fx = m.mv[l][b][0] & 3, fy = m.mv[l][b][1] & 3, cfx = m.mv[l][b][0] & 7, cfy = m.mv[l][b][1] & 7;
                alg_inter += (uint64_t)(w + 5 * (fx != 0)) * (h + 5 * (fy != 0)) +
                             2ull * (w / 2 + (cfx != 0)) * (h / 2 + (cfy != 0)) + 5;
                n_part_lists++;
            }
        }
        alg_inter += 384 + 2 * (uint64_t)(nmask + nvals) + 8;
    }
    // ---- queues
    pic->inter_list.clear();
    pic->intra_t.clear();
    pic->all_t.clear();
    for (uint32_t a = 0; a < n_mbs; a++)
        if (aux[a].kind & (AUX_INTER | AUX_PCM)) pic->inter_list.push_back(a);
    pic->mbaff = mbaff;
    if (mbaff) {  // pairs by t = x + 2 y, the top MB before the bottom one
        const int HP = H / 2;
        for (int t = 0; t <= (W - 1) + 2 * (HP - 1); t++)
            for (int y = 0; y < HP; y++) {
                const int xx = t - 2 * y;
                if (xx < 0) break;
                if (xx >= W) continue;
                const uint32_t pair = (uint32_t)(y * W + xx);
                pic->all_t.push_back(pair);
                for (uint32_t a = 2 * pair; a < 2 * pair + 2; a++)
                    if (aux[a].kind & AUX_INTRA) {
                        aux[a].ires |= (uint32_t)pic->intra_t.size();  // residual slot
                        pic->intra_t.push_back(a);
                    }
            }
        pic->n_pairs = (uint32_t)pic->all_t.size();
    } else {
        for (int t = 0; t <= (W - 1) + 2 * (H - 1); t++)
            for (int y = 0; y < H; y++) {
                int xx = t - 2 * y;
                if (xx < 0) break;
                if (xx >= W) continue;
                uint32_t a = (uint32_t)(y * W + xx);
                pic->all_t.push_back(a);
                if (aux[a].kind & AUX_INTRA) pic->intra_t.push_back(a);
            }
    }
    pic->n_inter = (uint32_t)pic->inter_list.size();
    pic->n_intra = (uint32_t)pic->intra_t.size();
    pic->n_mbs = mbaff ? pic->n_pairs : n_mbs;  // entries of all_t
    pic->alg[0] = alg_inter;
    pic->alg[1] = alg_intra;
    // deblocking: one read and one write of the picture, the bS inputs (motion records,
    // 8 bytes per MB and 5 per partition-list) and 4 bytes of non-zero / QP bits per MB
    pic->alg[2] = 2ull * 384 * n_mbs + 8ull * n_mbs + 5 * n_part_lists + 4ull * n_mbs;
    // ---- device block: mbs | aux | levels | weights | items | pic
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    size_t o_mbs = 0, o_aux = al(o_mbs + n_mbs * sizeof(h264g_mb)), o_lv = al(o_aux + n_mbs * sizeof(DevAux));
    size_t o_wt = al(o_lv + n_levels * sizeof(int16_t) + 16), o_it = al(o_wt + wt.size() * sizeof(DevWeight));
    pic->n_rows = mbaff ? 0u : (uint32_t)H;
    // intra MBs per row: offsets, then one entry per intra MB (the per-picture intra band
    // schedule; an MB's residual sits at its entry's index in P.ires).  The entry carries
    // what the band kernel would otherwise load from the neighbours' records before its
    // wait (ICOL_* in h264g_dev.h): the column, the neighbour availability (A, B, C, D in
    // bits 0..3; constrained_intra_pred drops the inter neighbours cip_hides names), which available
    // neighbours are intra MBs of this stage (C only when the MB reads its samples:
    // Intra_8x8, or Intra_4x4 block 5 in modes 3 / 7), and whether it has levels.
    std::vector<uint32_t> irow(H + 1, 0), icol;
    for (int y = 0; y < H && !mbaff; y++) {
        irow[y] = (uint32_t)icol.size();
        for (int xx = 0; xx < W; xx++) {
            const uint32_t a = (uint32_t)(y * W + xx);
            if (!(aux[a].kind & AUX_INTRA)) continue;
            uint32_t av = (xx > 0 ? 1u : 0u) | (y > 0 ? 2u : 0u) | (y > 0 && xx < W - 1 ? 4u : 0u) |
                          (xx > 0 && y > 0 ? 8u : 0u), dep = 0;
            const int64_t nb[4] = {(int64_t)a - 1, (int64_t)a - W, (int64_t)a - W + 1, (int64_t)a - W - 1};
            for (int k = 0; k < 4; k++) {
                if (!(av & (1u << k))) continue;
                const int c = mbs[nb[k]].mb_type;
                if (mbs[nb[k]].slice_num != mbs[a].slice_num) {  // another slice: not available (6.4.8)
                    av &= ~(1u << k);
                    continue;
                }
                if (d.constrained_intra_pred_flag && cip_hides(mbs[nb[k]])) av &= ~(1u << k);
                if (cls_intra(c)) dep |= 1u << k;
            }
            const h264g_mb &m = mbs[a];
            if (!(m.mb_type == H264G_MB_I8x8 ||
                  (m.mb_type == H264G_MB_I4x4 && (m.intra_pred_modes[5] == 3 || m.intra_pred_modes[5] == 7))))
                dep &= ~4u;
            aux[a].ires |= (uint32_t)icol.size();
            icol.push_back((uint32_t)xx | av << ICOL_AV_SHIFT | dep << ICOL_DEP_SHIFT |
                           ((aux[a].kind & AUX_RESID) ? ICOL_RES : 0u));
        }
    }
    irow[H] = (uint32_t)icol.size();
    size_t n_items = pic->n_inter + pic->n_intra + pic->all_t.size() + pic->n_rows + irow.size() + icol.size();
    size_t o_bs = al(o_it + n_items * sizeof(uint32_t)), o_dbq = al(o_bs + (size_t)n_mbs * 32);
    size_t o_ires = al(o_dbq + (size_t)n_mbs * 4), o_bsref = al(o_ires + (size_t)pic->n_intra * 768);
    size_t o_ls = al(o_bsref + bsref.size()), o_pic = al(o_ls + lsv.size() * sizeof(uint16_t));
    size_t total = al(o_pic + sizeof(DevPic));
    void *dm = fixed;
    if (fixed) {
        if (total > fixed_cap)
            return h264g_set_error(H264G_E_ARG, "picture needs %zu bytes of device staging, the fixed block holds %zu",
                                   total, fixed_cap);
    } else {
        std::lock_guard<std::mutex> lk(ctx->mu);  // the rest of the staging runs unlocked
        dm = pool_get(ctx, total);
        if (!dm) return h264g_set_error(H264G_E_NOMEM, "picture device memory (%zu bytes)", total);
        pic->dmem = dm;
        pic->dbytes = total;
    }
    char *base = (char *)dm;
    pic->d_mbs = (h264g_mb *)(base + o_mbs);
    pic->d_aux = (DevAux *)(base + o_aux);
    pic->d_levels = (int16_t *)(base + o_lv);
    pic->d_wt = (DevWeight *)(base + o_wt);
    pic->d_items = (uint32_t *)(base + o_it);
    pic->d_pic = (DevPic *)(base + o_pic);
    uint8_t *d_bs = (uint8_t *)(base + o_bs);
    uint32_t *d_dbq = (uint32_t *)(base + o_dbq);
    std::vector<uint32_t> items;
    items.reserve(n_items);
    items.insert(items.end(), pic->inter_list.begin(), pic->inter_list.end());
    items.insert(items.end(), pic->intra_t.begin(), pic->intra_t.end());
    items.insert(items.end(), pic->all_t.begin(), pic->all_t.end());
    for (uint32_t y = 0; y < pic->n_rows; y++) items.push_back(y);
    const size_t irow_at = items.size();
    items.insert(items.end(), irow.begin(), irow.end());
    items.insert(items.end(), icol.begin(), icol.end());
    DevPic &P = pic->hpic;
    std::memset(&P, 0, sizeof(P));
    P.mbs = pic->d_mbs; P.aux = pic->d_aux; P.levels = pic->d_levels; P.wt = pic->d_wt;
    P.dpb = st.dpb; P.slot_bytes = st.slot_bytes; P.bs = d_bs; P.dbq = d_dbq; P.dst = d.dst_slot;
    P.ires = (int16_t *)(base + o_ires);
    P.irow = pic->d_items + irow_at; P.n_intra = pic->n_intra;
    P.W = W; P.H = H; P.offA = d.filter_offset_a; P.offB = d.filter_offset_b;
    P.mbaff = mbaff ? 1 : 0;
    P.bsref = mbaff ? (const int8_t *)(base + o_bsref) : nullptr;
    use_set(st, 0, P);
    P.dbk_idc = d.disable_deblocking_filter_idc; P.cip = d.constrained_intra_pred_flag;
    P.dbk_stop = dbk_stop;
    P.multi_slice = multi_slice ? 1 : 0;
    P.ls = lsv.empty() ? nullptr : (const uint16_t *)(base + o_ls);
    HIPCHK(hipMemcpy(pic->d_mbs, mbs, n_mbs * sizeof(h264g_mb), hipMemcpyHostToDevice));
    if (!lsv.empty()) HIPCHK(hipMemcpy(base + o_ls, lsv.data(), lsv.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pic->d_aux, aux.data(), n_mbs * sizeof(DevAux), hipMemcpyHostToDevice));
    if (n_levels) HIPCHK(hipMemcpy(pic->d_levels, levels, n_levels * sizeof(int16_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pic->d_wt, wt.data(), wt.size() * sizeof(DevWeight), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pic->d_items, items.data(), items.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (mbaff) HIPCHK(hipMemcpy(base + o_bsref, bsref.data(), bsref.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pic->d_pic, &P, sizeof(DevPic), hipMemcpyHostToDevice));
    return 0;
}

// The launches of one picture or batch.  all = every MB (t order: it serves the bS
// pre-pass and the per-MB deblocking wavefront); rows = MB rows (row index major) for
// the cross-CU row schedule.
struct StageSet {
    const DevPic *pics;
    uint32_t n_pics;
    const uint32_t *inter, *intra, *all, *rows, *ibands, *dbands;
    uint32_t n_inter, n_intra, n_all, n_rows, n_ibands, n_dbands;
    bool dbk_off;
    bool scaled = false;  // a picture has scaling matrices: the inter stage takes the kernel that reads them
    // MBAFF frames: inter MBs, intra MBs for the per-MB wavefront, MB pairs for their
    // deblocking
    const uint32_t *minter = nullptr, *mintra = nullptr, *mpairs = nullptr;
    uint32_t n_minter = 0, n_mintra = 0, n_mpairs = 0;
};

// Parts of a stage set: inter / intra / bS, and deblocking.
constexpr int kFront = 1, kBack = 2;

// Launches the given parts on stream s.  ev: the run's 6 profiling events (inter, intra,
// bS, end of those; deblocking, end) or null.
int launch_stages(h264g_ctx *ctx, const StageSet &q, int lane, int parts, hipStream_t s, hipEvent_t *ev) {
    DevCtl *ctl = &ctx->d_ctl[4 * lane];
    int e = 0;
    if (parts & kFront) {
        if (ev) HIPCHK(hipEventRecord(ev[0], s));
        e = h264g_launch_inter(q.pics, q.inter, q.n_inter, q.intra, q.n_intra, 0, q.scaled ? 1 : 0, ctx->d_err, s);
        if (e) return h264g_set_error(H264G_E_DEVICE, "inter launch: %s", hipGetErrorString((hipError_t)e));
        // MBAFF frames: inter MBs and the residuals of their intra MBs (lists of their own)
        e = h264g_launch_inter(q.pics, q.minter, q.n_minter, q.mintra, q.n_mintra, 1, 0, ctx->d_err, s);
        if (e) return h264g_set_error(H264G_E_DEVICE, "inter launch (MBAFF): %s", hipGetErrorString((hipError_t)e));
        if (ev) HIPCHK(hipEventRecord(ev[1], s));
        if (q.n_intra)
            e = ctx->intra_schedule == H264G_INTRA_BANDS
                    ? h264g_launch_intra_band(q.pics, q.ibands, q.n_ibands, &ctl[0], &ctl[2].ticket, ctx->intra_waves,
                                              ctx->d_err, s)
                    : h264g_launch_intra(q.pics, q.intra, q.n_intra, &ctl[0], &ctl[2].ticket, ctx->d_err, ctx->waves, 0,
                                         s);
        if (e) return h264g_set_error(H264G_E_DEVICE, "intra launch: %s", hipGetErrorString((hipError_t)e));
        if (q.n_mintra)
            e = h264g_launch_intra(q.pics, q.mintra, q.n_mintra, &ctl[0], &ctl[2].ticket, ctx->d_err, ctx->waves, 1, s);
        if (e) return h264g_set_error(H264G_E_DEVICE, "intra launch (MBAFF): %s", hipGetErrorString((hipError_t)e));
        if (ev) HIPCHK(hipEventRecord(ev[2], s));
        if (!q.dbk_off) {
            e = h264g_launch_bs(q.pics, q.all, q.n_all, s);
            if (e) return h264g_set_error(H264G_E_DEVICE, "bS launch: %s", hipGetErrorString((hipError_t)e));
        }
        if (ev) HIPCHK(hipEventRecord(ev[3], s));
    }
    if (parts & kBack) {
        if (ev) HIPCHK(hipEventRecord(ev[4], s));
        if (!q.dbk_off) {
            uint32_t *epoch = &ctl[3].ticket;
            switch (ctx->dbk_schedule) {
            case H264G_DEBLOCK_BANDS:
                e = h264g_launch_deblock_band(q.pics, q.dbands, q.n_dbands, &ctl[1], epoch, ctx->dbk_waves,
                                              ctx->d_err, s);
                break;
            case H264G_DEBLOCK_MB_WAVEFRONT:
                e = h264g_launch_deblock(q.pics, q.all, q.n_all, &ctl[1], epoch, ctx->d_err, ctx->waves, s);
                break;
            default:
                e = h264g_launch_deblock_rows(q.pics, q.rows, q.n_rows, &ctl[1], epoch, ctx->d_err, ctx->waves, s);
            }
            if (e) return h264g_set_error(H264G_E_DEVICE, "deblock launch: %s", hipGetErrorString((hipError_t)e));
            e = h264g_launch_deblock_mbaff(q.pics, q.mpairs, q.n_mpairs, &ctl[1], epoch, ctx->d_err, ctx->waves, s);
            if (e) return h264g_set_error(H264G_E_DEVICE, "deblock launch (MBAFF): %s", hipGetErrorString((hipError_t)e));
        }
        if (ev) HIPCHK(hipEventRecord(ev[5], s));
    }
    return 0;
}

hipEvent_t *profile_events(h264g_ctx *ctx) {
    if (!ctx->profile) return nullptr;
    while (ctx->ev.size() < ctx->ev_used + 6) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->ev.push_back(e);
    }
    hipEvent_t *ev = &ctx->ev[ctx->ev_used];
    ctx->ev_used += 6;
    return ev;
}

bool conflicts(const DbkRec &r, const std::vector<SlotUse> &uses) {
    for (const auto &d : r.dst)
        for (const SlotUse &u : uses)
            if (u.stream == d.first && (u.dst == d.second || std::find(u.refs.begin(), u.refs.end(), d.second) != u.refs.end()))
                return true;
    return false;
}

// Stages of `q` (pictures `uses`) on `lane`.  sb: the batch whose captured graphs to replay
// (captured on first use), or null for direct launches.  Order: the lane waits for the
// newest deblocking launch in flight that writes a slot these pictures write or read
// (deblocking launches on the lane's deblocking stream complete in order, so that one
// covers the older ones); inter / intra / bS on the lane; deblocking on the deblocking
// stream after them, with the band schedule in a two-lane context (the other schedules
// share per-MB progress words with the intra stage, and a one-lane context is the
// strictly sequential shape: there deblocking stays on the lane).
int enqueue(h264g_ctx *c, const StageSet &q, int lane, const std::vector<SlotUse> &uses, Batch *sb,
            bool events = true) {
    hipStream_t s = lane_stream(c, lane), ds = c->dstream[lane];
    for (int k = 1; k <= kDbkHist; k++) {
        const DbkRec &r = c->hist[lane][(c->hist_next[lane] - k + kDbkHist) % kDbkHist];
        if (r.live && conflicts(r, uses)) {
            HIPCHK(hipStreamWaitEvent(s, r.ev, 0));
            break;
        }
    }
    // one-lane contexts (lane_split_mbs < 0) keep the strictly sequential shape
    const bool split = c->dbk_schedule == H264G_DEBLOCK_BANDS && !q.dbk_off && c->split_mbs > 0;
    hipEvent_t *ev = events ? profile_events(c) : nullptr;
    const bool graphs = sb && !c->profile && !c->no_graphs;
    // capture of the launches `parts` on stream st into *g (kernel nodes only; every launch
    // argument, including the wavefront epochs, which live in device memory, is the same
    // on each replay)
    auto capture = [&](int parts, hipStream_t st, hipGraphExec_t *g) -> int {
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = launch_stages(c, q, lane, parts, st, nullptr);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(st, &gr);
        if (rc) { if (gr) (void)hipGraphDestroy(gr); return rc; }
        if (e != hipSuccess) return h264g_set_error(H264G_E_DEVICE, "stream capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(g, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) { *g = nullptr; return h264g_set_error(H264G_E_DEVICE, "graph instantiate: %s", hipGetErrorString(e)); }
        return 0;
    };
    const int first = split ? kFront : (kFront | kBack);
    if (graphs) {
        if (!sb->graph) { int rc = capture(first, s, &sb->graph); if (rc) return rc; }
        HIPCHK(hipGraphLaunch(sb->graph, s));
    } else {
        int rc = launch_stages(c, q, lane, first, s, ev);
        if (rc) return rc;
    }
    if (!split) return 0;
    HIPCHK(hipEventRecord(c->front[lane], s));
    HIPCHK(hipStreamWaitEvent(ds, c->front[lane], 0));
    if (graphs) {
        if (!sb->graph_dbk) { int rc = capture(kBack, ds, &sb->graph_dbk); if (rc) return rc; }
        HIPCHK(hipGraphLaunch(sb->graph_dbk, ds));
    } else {
        int rc = launch_stages(c, q, lane, kBack, ds, ev);
        if (rc) return rc;
    }
    DbkRec &r = c->hist[lane][c->hist_next[lane]];
    if (r.live) HIPCHK(hipStreamWaitEvent(s, r.ev, 0));  // leaves the history: order the lane after it
    if (!r.ev) HIPCHK(hipEventCreateWithFlags(&r.ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(r.ev, ds));
    r.dst.clear();
    for (const SlotUse &u : uses) r.dst.push_back({u.stream, u.dst});
    r.live = true;
    c->hist_next[lane] = (c->hist_next[lane] + 1) % kDbkHist;
    return 0;
}

// Caller holds ctx->mu.  A batch with a graph is destroyed only when idle.
int free_batch(h264g_ctx *c, h264g_batch *b) {
    if (!b) return 0;
    if (b->sub[0].graph || b->sub[1].graph || b->sub[0].graph_dbk || b->sub[1].graph_dbk) (void)sync_lanes(c);
    for (Batch &sb : b->sub) {
        if (sb.graph) (void)hipGraphExecDestroy(sb.graph);
        if (sb.graph_dbk) (void)hipGraphExecDestroy(sb.graph_dbk);
        pool_put(c, sb.dmem, sb.dbytes);
    }
    delete b;
    return 0;
}

}  // namespace

extern "C" {

const char *h264g_last_error(void) { return g_err; }

int32_t h264g_device_describe(int32_t device, char *out, uint32_t len) {
    if (!out || len == 0) return h264g_set_error(H264G_E_ARG, "NULL argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return h264g_set_error(H264G_E_NODEV, "device %d", device);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    snprintf(out, len, "%s|%s|%d CUs|%d kHz", prop.gcnArchName, prop.name, prop.multiProcessorCount, prop.clockRate);
    return H264G_OK;
}

// Progress words start at 0, so the first wavefront launch uses epoch 1.  Intra and
// deblocking launches count their epochs separately (they may run at the same time on a
// lane and its deblocking stream); the deblocking count starts half the range away, so
// the two never meet in the per-MB progress words the per-MB schedules share.
static const uint32_t kFirstEpoch = 1, kFirstDbkEpoch = 0x80000001u;

int32_t h264g_create(const h264g_config *cfg, h264g_ctx **out) {
    if (!out) return h264g_set_error(H264G_E_ARG, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return h264g_set_error(H264G_E_NODEV, "no HIP device");
    int dev = cfg ? cfg->device : 0;
    if (dev < 0 || dev >= ndev) return h264g_set_error(H264G_E_NODEV, "device %d of %d", dev, ndev);
    HIPCHK(hipSetDevice(dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return h264g_set_error(H264G_E_NODEV, "device %d is %s, the kernels are built for gfx950", dev, prop.gcnArchName);
    h264g_ctx *c = new (std::nothrow) h264g_ctx();
    if (!c) return h264g_set_error(H264G_E_NOMEM, "ctx");
    c->device = dev;
    c->budget = cfg ? cfg->dpb_budget_bytes : 0;
    c->profile = cfg && cfg->profile;
    c->no_graphs = cfg && cfg->no_graphs;
    c->dbk_schedule = cfg ? cfg->deblock_schedule : H264G_DEBLOCK_BANDS;
    c->dbk_waves = (cfg && cfg->deblock_waves > 0) ? cfg->deblock_waves : 8;
    c->intra_schedule = cfg ? cfg->intra_schedule : H264G_INTRA_BANDS;
    c->intra_waves = (cfg && cfg->intra_waves > 0) ? cfg->intra_waves : 8;
    if (c->intra_schedule < 0 || c->intra_schedule > 1 || c->intra_waves > 8) {
        h264g_destroy(c);
        return h264g_set_error(H264G_E_ARG, "intra_schedule %d / intra_waves %d", cfg->intra_schedule, cfg->intra_waves);
    }
    if (c->dbk_schedule < 0 || c->dbk_schedule > 2 || c->dbk_waves < 1 || c->dbk_waves > 24) {
        h264g_destroy(c);
        return h264g_set_error(H264G_E_ARG, "deblock_schedule %d / deblock_waves %d", cfg->deblock_schedule, cfg->deblock_waves);
    }
    c->waves = (cfg && cfg->waves > 0) ? cfg->waves : prop.multiProcessorCount * 8;
    c->split_mbs = !cfg || cfg->lane_split_mbs == 0 ? 16384 : (cfg->lane_split_mbs < 0 ? 0 : cfg->lane_split_mbs);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->dstream[0], hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->dstream[1], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->front[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->front[1], hipEventDisableTiming) != hipSuccess ||
        hipMalloc(&c->d_err, sizeof(uint32_t)) != hipSuccess || hipMalloc(&c->d_ctl, 8 * sizeof(DevCtl)) != hipSuccess ||
        hipMemset(c->d_err, 0, sizeof(uint32_t)) != hipSuccess || hipMemset(c->d_ctl, 0, 8 * sizeof(DevCtl)) != hipSuccess ||
        hipMemcpy(&c->d_ctl[2].ticket, &kFirstEpoch, sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(&c->d_ctl[3].ticket, &kFirstDbkEpoch, sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(&c->d_ctl[6].ticket, &kFirstEpoch, sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(&c->d_ctl[7].ticket, &kFirstDbkEpoch, sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
        h264g_destroy(c);
        return h264g_set_error(H264G_E_DEVICE, "context allocation failed");
    }
    // events of the deblocking history, so that enqueuing creates none
    for (auto &lane : c->hist)
        for (auto &r : lane)
            if (hipEventCreateWithFlags(&r.ev, hipEventDisableTiming) != hipSuccess) {
                h264g_destroy(c);
                return h264g_set_error(H264G_E_DEVICE, "context events");
            }
    *out = c;
    return H264G_OK;
}

int32_t h264g_destroy(h264g_ctx *c) {
    if (!c) return H264G_OK;
    (void)hipSetDevice(c->device);
    for (int i = 0; i < 4; i++)
        if (any_stream(c, i)) (void)hipStreamSynchronize(any_stream(c, i));
    for (auto &b : c->pool) {
        for (auto ev : b.ready) (void)hipEventDestroy(ev);
        (void)hipFree(b.ptr);
    }
    c->pool.clear();
    for (auto &s : c->streams) {
        if (s.dpb) (void)hipFree(s.dpb);
        if (s.prog) (void)hipFree(s.prog);
    }
    for (auto &e : c->ev) if (e) (void)hipEventDestroy(e);
    if (c->d_err) (void)hipFree(c->d_err);
    if (c->d_ctl) (void)hipFree(c->d_ctl);
    for (int l = 0; l < 2; l++) {
        for (auto &r : c->hist[l]) if (r.ev) (void)hipEventDestroy(r.ev);
        if (c->front[l]) (void)hipEventDestroy(c->front[l]);
    }
    for (int i = 0; i < 4; i++)
        if (any_stream(c, i)) (void)hipStreamDestroy(any_stream(c, i));
    delete c;
    return H264G_OK;
}

int32_t h264g_stream_open(h264g_ctx *c, const h264g_stream_desc *d, uint32_t *id) {
    if (!c || !d || !id) return h264g_set_error(H264G_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (d->pic_width_in_mbs <= 0 || d->frame_height_in_mbs <= 0 || d->pic_width_in_mbs * d->frame_height_in_mbs >= (1 << 20))
        return h264g_set_error(H264G_E_ARG, "picture size");
    if (d->chroma_array_type != 1 || d->bit_depth != 8) return h264g_set_error(H264G_E_ARG, "only 8-bit 4:2:0 is supported");
    if (d->dpb_slots < 1 || d->dpb_slots > 17) return h264g_set_error(H264G_E_ARG, "dpb_slots");
    if (d->frame_height_in_mbs > 1024) return h264g_set_error(H264G_E_ARG, "frame_height_in_mbs > 1024");
    if (d->pic_width_in_mbs > (int)ICOL_X_MASK + 1) return h264g_set_error(H264G_E_ARG, "pic_width_in_mbs > 2048");
    if (d->frame_mbs_only && d->mbaff) return h264g_set_error(H264G_E_ARG, "mbaff needs frame_mbs_only = 0");
    if (!d->frame_mbs_only && !d->mbaff) return h264g_set_error(H264G_E_ARG, "field pictures are not supported");
    if (d->mbaff && d->frame_height_in_mbs % 2) return h264g_set_error(H264G_E_ARG, "MBAFF frames need an even height");
    Stream s;
    s.desc = *d;
    s.W = d->pic_width_in_mbs;
    s.H = d->frame_height_in_mbs;
    s.mbaff = d->mbaff != 0;
    // A slot is Y, Cb, Cr; MBAFF slots end in a zeroed chroma row: field MBs read up to one
    // chroma row past the Cr plane, as the reference does past its frame buffer (kernels.hip
    // luma_list), and find zeros there as the reference does
    s.slot_bytes = (uint64_t)s.W * s.H * 384 + (s.mbaff ? ((uint64_t)s.W * 8 + 255) / 256 * 256 : 0);
    s.lane = c->split_mbs > 0 && s.W * s.H >= c->split_mbs ? 1 : 0;
    // DPB slots, and the progress sets: per-MB progress words, per-row progress words,
    // the strips between row bands (deblocking: 256 B per MB column and band; intra: 64 B)
    // and the MB pair progress words of MBAFF streams
    auto al = [](uint64_t v) { return (v + 255) & ~(uint64_t)255; };
    s.o_rowprog = al((uint64_t)s.W * s.H * 4);
    s.o_gstrip = al(s.o_rowprog + (uint64_t)s.H * 8);
    s.o_istrip = s.o_gstrip + (uint64_t)((s.H + c->dbk_waves - 1) / c->dbk_waves) * s.W * 256;
    s.o_pair = al(s.o_istrip + (uint64_t)((s.H + c->intra_waves - 1) / c->intra_waves) * s.W * 64);
    s.set_bytes = al(s.o_pair + (s.mbaff ? (uint64_t)s.W * s.H / 2 * 4 : 0));
    const size_t dpb_bytes = s.slot_bytes * d->dpb_slots, prog_bytes = s.set_bytes * H264G_MAX_STREAM_PICS;
    const uint64_t need = dpb_bytes + prog_bytes;
    if (c->budget && c->used + need > c->budget)
        return h264g_set_error(H264G_E_BUDGET, "DPB budget: %llu + %llu > %llu", (unsigned long long)c->used,
                               (unsigned long long)need, (unsigned long long)c->budget);
    HIPCHK(hipSetDevice(c->device));
    // Blocks come from the event-guarded pool (a closed stream's blocks return there), and
    // the clears run on the stream's lane ahead of its first picture, so opening a stream
    // neither allocates in steady state nor waits for the GPU.
    s.dpb = (uint8_t *)pool_get(c, dpb_bytes);
    s.prog = (uint8_t *)pool_get(c, prog_bytes);
    if (!s.dpb || !s.prog) {
        pool_put(c, s.dpb, dpb_bytes);
        pool_put(c, s.prog, prog_bytes);
        return h264g_set_error(H264G_E_NOMEM, "stream device memory");
    }
    hipStream_t ls = lane_stream(c, s.lane);
    HIPCHK(hipMemsetAsync(s.prog, 0, prog_bytes, ls));  // epoch 0 never matches a launch
    HIPCHK(hipMemsetAsync(s.dpb, 0, dpb_bytes, ls));
    s.strip_dbk_rows = c->dbk_waves;
    s.strip_intra_rows = c->intra_waves;
    s.bytes = need;
    s.open = true;
    s.gen = ++c->opened;
    c->used += need;
    uint32_t i = 0;
    for (; i < c->streams.size(); i++) if (!c->streams[i].open) break;
    if (i == c->streams.size()) c->streams.push_back(s);
    else c->streams[i] = s;
    *id = i;
    return H264G_OK;
}

int32_t h264g_set_bands(h264g_ctx *c, int32_t deblock_rows, int32_t intra_rows) {
    if (!c) return h264g_set_error(H264G_E_ARG, "ctx");
    std::lock_guard<std::mutex> lk(c->mu);
    const int d = deblock_rows > 0 ? deblock_rows : c->dbk_waves, i = intra_rows > 0 ? intra_rows : c->intra_waves;
    if (d > 24 || i > 8) return h264g_set_error(H264G_E_ARG, "band heights %d / %d", d, i);
    for (uint32_t k = 0; k < c->streams.size(); k++) {
        const Stream &s = c->streams[k];
        if (s.open && (d < s.strip_dbk_rows || i < s.strip_intra_rows))
            return h264g_set_error(H264G_E_STATE, "stream %u has strips for bands of at least %d / %d rows", k,
                                   s.strip_dbk_rows, s.strip_intra_rows);
    }
    c->dbk_waves = d;
    c->intra_waves = i;
    return H264G_OK;
}

int32_t h264g_stream_close(h264g_ctx *c, uint32_t id) {
    if (!c) return h264g_set_error(H264G_E_ARG, "ctx");
    std::lock_guard<std::mutex> lk(c->mu);
    if (id >= c->streams.size() || !c->streams[id].open) return h264g_set_error(H264G_E_ARG, "stream %u", id);
    (void)hipSetDevice(c->device);
    Stream &s = c->streams[id];
    // back to the pool, reusable once the work enqueued so far has completed (no wait)
    pool_put(c, s.dpb, s.slot_bytes * s.desc.dpb_slots);
    pool_put(c, s.prog, s.set_bytes * H264G_MAX_STREAM_PICS);
    c->used -= s.bytes;
    s = Stream();
    return H264G_OK;
}

int32_t h264g_picture_begin(h264g_ctx *c, uint32_t sid, const h264g_pic_desc *d, const h264g_mb *mbs, uint32_t n_mbs,
                            const int16_t *levels, uint64_t n_levels, h264g_pic **out) {
    if (!c || !d || !mbs || !out || (!levels && n_levels)) return h264g_set_error(H264G_E_ARG, "NULL argument");
    *out = nullptr;
    // Only the stream lookup and the device block take the context lock: the host work
    // of staging (validation, per-MB records, work lists) and the uploads of several
    // pictures can run on several threads at once.
    Stream st;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (sid >= c->streams.size() || !c->streams[sid].open) return h264g_set_error(H264G_E_ARG, "stream %u", sid);
        st = c->streams[sid];
    }
    HIPCHK(hipSetDevice(c->device));
    h264g_pic *p = new (std::nothrow) h264g_pic();
    if (!p) return h264g_set_error(H264G_E_NOMEM, "pic");
    p->ctx = c;
    p->stream = sid;
    p->stream_gen = st.gen;
    p->desc = *d;
    p->W = st.W;
    p->H = st.H;
    int rc = stage_picture(c, st, *d, mbs, n_mbs, levels, n_levels, p);
    if (rc) {
        std::lock_guard<std::mutex> lk(c->mu);
        pool_put(c, p->dmem, p->dbytes);
        delete p;
        return rc;
    }
    *out = p;
    return H264G_OK;
}

namespace {
// The launches of one staged picture (caller holds ctx->mu).  events: profiling contexts
// record stage events (created on demand).
int enqueue_picture(h264g_ctx *c, const h264g_pic *p, bool events = true) {
    const uint32_t *it = p->d_items;
    StageSet q;
    q.pics = p->d_pic;
    q.n_pics = 1;
    q.inter = it;
    q.n_inter = p->n_inter;
    q.intra = it + p->n_inter;
    q.n_intra = p->n_intra;
    q.all = q.intra + p->n_intra;
    q.n_all = p->n_mbs;
    q.rows = q.all + p->n_mbs;
    q.n_rows = p->n_rows;
    q.ibands = q.dbands = q.rows;  // band b of picture 0 is item b
    q.n_ibands = (uint32_t)((p->H + c->intra_waves - 1) / c->intra_waves);
    q.n_dbands = (uint32_t)((p->H + c->dbk_waves - 1) / c->dbk_waves);
    q.dbk_off = p->desc.disable_deblocking_filter_idc == 1;
    q.scaled = p->hpic.ls != nullptr;
    if (p->mbaff) {  // items: inter MBs, intra MBs (pair t order), MB pairs (t order)
        q.minter = q.inter;
        q.n_minter = q.n_inter;
        q.n_inter = 0;
        q.mintra = q.intra;
        q.n_mintra = p->n_intra;
        q.mpairs = q.all;
        q.n_mpairs = q.dbk_off ? 0u : p->n_pairs;
        q.n_intra = q.n_all = q.n_rows = q.n_ibands = q.n_dbands = 0;
    }
    return enqueue(c, q, c->streams[p->stream].lane, {slot_use(p->stream, p->desc)}, nullptr, events);
}

// State of the allocation-free entry point per device: the lock that serialises its calls
// and the events recorded after a call's launches (on the lane, then on its deblocking
// stream), which the next call on the device waits for before it overwrites the static
// scratch array.  The events are created by the first call on the device.
constexpr int kStaticDevices = 64;
struct StaticState {
    std::mutex mu;
    hipEvent_t done[2] = {nullptr, nullptr};
};
StaticState g_static[kStaticDevices];
}  // namespace

int32_t h264g_picture_decode_static(h264g_ctx *c, uint32_t sid, const h264g_pic_desc *d, const h264g_mb *mbs,
                                    uint32_t n_mbs, const int16_t *levels, uint64_t n_levels) {
    if (!c || !d || !mbs || (!levels && n_levels)) return h264g_set_error(H264G_E_ARG, "NULL argument");
    if (c->device >= kStaticDevices) return h264g_set_error(H264G_E_NODEV, "device %d", c->device);
    Stream st;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (sid >= c->streams.size() || !c->streams[sid].open) return h264g_set_error(H264G_E_ARG, "stream %u", sid);
        st = c->streams[sid];
    }
    HIPCHK(hipSetDevice(c->device));
    StaticState &ss = g_static[c->device];
    std::lock_guard<std::mutex> lk(ss.mu);
    void *base = nullptr;
    if (int e = h264g_scratch_address(c->device, &base))
        return h264g_set_error(H264G_E_DEVICE, "static scratch address: %s", hipGetErrorString((hipError_t)e));
    for (hipEvent_t &ev : ss.done) {
        if (!ev) HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        else HIPCHK(hipEventSynchronize(ev));  // the previous call's launches have read the array
    }
    h264g_pic p;
    p.ctx = c;
    p.stream = sid;
    p.stream_gen = st.gen;
    p.desc = *d;
    p.W = st.W;
    p.H = st.H;
    if (int rc = stage_picture(c, st, *d, mbs, n_mbs, levels, n_levels, &p, base, H264G_STATIC_SCRATCH_BYTES)) return rc;
    std::lock_guard<std::mutex> lk2(c->mu);
    if (int rc = check_stream(c, sid, st.gen)) return rc;
    if (int rc = enqueue_picture(c, &p, false)) return rc;
    const int lane = c->streams[sid].lane;
    HIPCHK(hipEventRecord(ss.done[0], lane_stream(c, lane)));
    HIPCHK(hipEventRecord(ss.done[1], c->dstream[lane]));
    return H264G_OK;
}

int32_t h264g_picture_submit(h264g_ctx *c, h264g_pic *p) {
    if (!c || !p || p->ctx != c) return h264g_set_error(H264G_E_ARG, "bad picture");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = check_stream(c, p->stream, p->stream_gen)) return rc;
    HIPCHK(hipSetDevice(c->device));
    return enqueue_picture(c, p);
}

int32_t h264g_picture_release(h264g_ctx *c, h264g_pic *p) {
    if (!p) return H264G_OK;
    if (!c || p->ctx != c) return h264g_set_error(H264G_E_ARG, "bad picture");
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    pool_put(c, p->dmem, p->dbytes);  // reused only after the launches enqueued so far
    delete p;
    return H264G_OK;
}

int32_t h264g_picture_alg_bytes(const h264g_pic *p, uint64_t out[3]) {
    if (!p || !out) return h264g_set_error(H264G_E_ARG, "NULL argument");
    for (int i = 0; i < 3; i++) out[i] = p->alg[i];
    return H264G_OK;
}

namespace {
// Device lists of the pictures of one lane of a batch (caller holds ctx->mu).
int build_sub(h264g_ctx *c, const std::vector<const h264g_pic *> &pics, Batch *b) {
    const uint32_t n = (uint32_t)pics.size();
    std::vector<DevPic> hp(n);
    std::vector<uint32_t> inter, intra, dbk, minter;
    int maxt = 0;
    b->uses.clear();
    b->streams.clear();
    for (uint32_t i = 0; i < n; i++) {
        const h264g_pic *p = pics[i];
        hp[i] = p->hpic;
        // the k-th picture of a stream in the batch (a stream's pictures share its lane)
        // waits on and publishes to progress set k
        int k = 0;
        for (uint32_t j = 0; j < i; j++) k += pics[j]->stream == p->stream;
        use_set(c->streams[p->stream], k, hp[i]);
        b->uses.push_back(slot_use(p->stream, p->desc));
        b->streams.push_back({p->stream, p->stream_gen});
        for (uint32_t a : p->inter_list) (p->mbaff ? minter : inter).push_back((i << WI_PIC_SHIFT) | a);
        maxt = std::max(maxt, (p->W - 1) + 2 * (p->H - 1));
    }
    // the frame pictures' inter MBs, then the MBAFF frames' (each set is one launch)
    const uint32_t n_frame_inter = (uint32_t)inter.size();
    inter.insert(inter.end(), minter.begin(), minter.end());
    // Wavefront queues of all pictures merged by t = x + 2y (counting sort): every
    // dependency of an entry has a smaller t and therefore an earlier ticket.  MBAFF frames
    // (mbaff = true) go by the t of their MB pairs, the top MB before the bottom one.
    auto merge = [&](bool intra_only, bool mbaff, std::vector<uint32_t> &outq) {
        std::vector<std::vector<uint32_t>> byt(maxt + 1);
        for (uint32_t i = 0; i < n; i++) {
            const h264g_pic *p = pics[i];
            if (p->mbaff != mbaff || (!intra_only && p->desc.disable_deblocking_filter_idc == 1)) continue;
            const std::vector<uint32_t> &q = intra_only ? p->intra_t : p->all_t;
            for (uint32_t a : q) {
                const uint32_t xy = mbaff && intra_only ? a >> 1 : a;  // MB pairs: all_t holds pairs
                int t = (int)(xy % p->W) + 2 * (int)(xy / p->W);
                byt[t].push_back((i << WI_PIC_SHIFT) | a);
            }
        }
        for (auto &v : byt) outq.insert(outq.end(), v.begin(), v.end());
    };
    merge(true, false, intra);
    std::vector<uint32_t> mintra, mpairs;
    merge(true, true, mintra);
    merge(false, true, mpairs);
    bool any_dbk = false;
    for (uint32_t i = 0; i < n; i++) any_dbk |= !pics[i]->mbaff && pics[i]->desc.disable_deblocking_filter_idc != 1;
    std::vector<uint32_t> all, rows;
    if (any_dbk) merge(false, false, all);
    int maxh = 0;
    for (uint32_t i = 0; i < n; i++) maxh = std::max(maxh, pics[i]->H);
    // bands of R rows of all pictures, band index major: the band above always comes first
    auto bands = [&](int R, bool intra, std::vector<uint32_t> &out) {
        for (int b = 0; b * R < maxh; b++)
            for (uint32_t i = 0; i < n; i++) {
                const h264g_pic *p = pics[i];
                if (!p->mbaff && b * R < p->H && (intra ? p->n_intra != 0 : p->desc.disable_deblocking_filter_idc != 1))
                    out.push_back((i << WI_PIC_SHIFT) | (uint32_t)b);
            }
    };
    std::vector<uint32_t> ibands, dbands;
    if (!intra.empty()) bands(c->intra_waves, true, ibands);
    if (any_dbk) bands(c->dbk_waves, false, dbands);
    if (any_dbk && c->dbk_schedule == H264G_DEBLOCK_MB_ROWS) {
        // MB rows of all pictures, row index major: the row above always comes first.
        for (int y = 0; y < maxh; y++)
            for (uint32_t i = 0; i < n; i++)
                if (!pics[i]->mbaff && y < pics[i]->H) rows.push_back((i << WI_PIC_SHIFT) | (uint32_t)y);
    }
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    size_t o_p = 0, o_i = al(n * sizeof(DevPic)), o_a = al(o_i + inter.size() * 4), o_l = al(o_a + intra.size() * 4);
    size_t o_r = al(o_l + all.size() * 4), o_ib = al(o_r + rows.size() * 4), o_db = al(o_ib + ibands.size() * 4);
    size_t o_mi = al(o_db + dbands.size() * 4), o_mp = al(o_mi + mintra.size() * 4);
    size_t total = al(o_mp + mpairs.size() * 4 + 4);
    b->dmem = pool_get(c, total);
    b->dbytes = total;
    if (!b->dmem) return h264g_set_error(H264G_E_NOMEM, "batch device memory");
    char *base = (char *)b->dmem;
    b->d_pics = (DevPic *)(base + o_p);
    b->n_pics = n;
    b->d_inter = (uint32_t *)(base + o_i);
    b->d_intra = (uint32_t *)(base + o_a);
    b->d_all = (uint32_t *)(base + o_l);
    b->d_rows = (uint32_t *)(base + o_r);
    b->d_ibands = (uint32_t *)(base + o_ib);
    b->d_dbands = (uint32_t *)(base + o_db);
    b->n_inter = n_frame_inter;
    b->d_minter = b->d_inter + n_frame_inter;
    b->n_minter = (uint32_t)minter.size();
    b->n_intra = (uint32_t)intra.size();
    b->n_all = (uint32_t)all.size();
    b->n_rows = (uint32_t)rows.size();
    b->n_ibands = (uint32_t)ibands.size();
    b->n_dbands = (uint32_t)dbands.size();
    b->d_mintra = (uint32_t *)(base + o_mi);
    b->d_mpairs = (uint32_t *)(base + o_mp);
    b->n_mintra = (uint32_t)mintra.size();
    b->n_mpairs = (uint32_t)mpairs.size();
    b->dbk_off = !any_dbk && mpairs.empty();
    for (uint32_t i = 0; i < n; i++) b->scaled |= hp[i].ls != nullptr;
    if (hipMemcpy(b->d_pics, hp.data(), n * sizeof(DevPic), hipMemcpyHostToDevice) != hipSuccess ||
        (!inter.empty() && hipMemcpy(b->d_inter, inter.data(), inter.size() * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (!intra.empty() && hipMemcpy(b->d_intra, intra.data(), intra.size() * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (!all.empty() && hipMemcpy(b->d_all, all.data(), all.size() * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (!rows.empty() && hipMemcpy(b->d_rows, rows.data(), rows.size() * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (!ibands.empty() && hipMemcpy(b->d_ibands, ibands.data(), ibands.size() * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (!dbands.empty() && hipMemcpy(b->d_dbands, dbands.data(), dbands.size() * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (!mintra.empty() && hipMemcpy(b->d_mintra, mintra.data(), mintra.size() * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (!mpairs.empty() && hipMemcpy(b->d_mpairs, mpairs.data(), mpairs.size() * 4, hipMemcpyHostToDevice) != hipSuccess))
        return h264g_set_error(H264G_E_DEVICE, "batch upload");
    return H264G_OK;
}
}  // namespace

int32_t h264g_batch_create(h264g_ctx *c, h264g_pic *const *pics, uint32_t n, h264g_batch **out) {
    if (!c || !pics || !out || n == 0) return h264g_set_error(H264G_E_ARG, "NULL argument");
    *out = nullptr;
    if (n >= (1u << (32 - WI_PIC_SHIFT))) return h264g_set_error(H264G_E_ARG, "too many pictures in one batch");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    std::vector<const h264g_pic *> lane[2];
    for (uint32_t i = 0; i < n; i++) {
        const h264g_pic *p = pics[i];
        if (!p || p->ctx != c) return h264g_set_error(H264G_E_ARG, "picture %u", i);
        if (int rc = check_stream(c, p->stream, p->stream_gen)) return rc;
        // pictures of one stream: at most H264G_MAX_STREAM_PICS (one progress set each),
        // none writing a slot another writes or reads
        const SlotUse u = slot_use(p->stream, p->desc);
        int same = 0;
        for (uint32_t j = 0; j < i; j++) {
            if (!pics[j] || pics[j]->stream != p->stream) continue;
            const SlotUse v = slot_use(pics[j]->stream, pics[j]->desc);
            if (u.dst == v.dst || std::find(u.refs.begin(), u.refs.end(), v.dst) != u.refs.end() ||
                std::find(v.refs.begin(), v.refs.end(), u.dst) != v.refs.end())
                return h264g_set_error(H264G_E_ARG, "pictures %u and %u of stream %u in one batch depend on each other "
                                       "(slots %d and %d)", j, i, p->stream, v.dst, u.dst);
            if (++same >= H264G_MAX_STREAM_PICS)
                return h264g_set_error(H264G_E_ARG, "more than %d pictures of stream %u in one batch",
                                       H264G_MAX_STREAM_PICS, p->stream);
        }
        lane[c->streams[p->stream].lane].push_back(p);
    }
    h264g_batch *b = new (std::nothrow) h264g_batch();
    if (!b) return h264g_set_error(H264G_E_NOMEM, "batch");
    b->dbk_rows = c->dbk_waves;
    b->intra_rows = c->intra_waves;
    for (int l = 0; l < 2; l++) {
        if (lane[l].empty()) continue;
        int rc = build_sub(c, lane[l], &b->sub[l]);
        if (rc) {
            free_batch(c, b);
            return rc;
        }
    }
    *out = b;
    return H264G_OK;
}

namespace {
StageSet batch_stages(const Batch *b) {
    StageSet q;
    q.pics = b->d_pics;
    q.n_pics = b->n_pics;
    q.inter = b->d_inter;
    q.n_inter = b->n_inter;
    q.intra = b->d_intra;
    q.n_intra = b->n_intra;
    q.all = b->d_all;
    q.n_all = b->n_all;
    q.rows = b->d_rows;
    q.n_rows = b->n_rows;
    q.ibands = b->d_ibands;
    q.n_ibands = b->n_ibands;
    q.dbands = b->d_dbands;
    q.n_dbands = b->n_dbands;
    q.dbk_off = b->dbk_off;
    q.scaled = b->scaled;
    q.minter = b->d_minter;
    q.n_minter = b->n_minter;
    q.mintra = b->d_mintra;
    q.n_mintra = b->n_mintra;
    q.mpairs = b->d_mpairs;
    q.n_mpairs = b->n_mpairs;
    return q;
}
}  // namespace

int32_t h264g_batch_run(h264g_ctx *c, h264g_batch *b) {
    if (!c || !b) return h264g_set_error(H264G_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (b->dbk_rows != c->dbk_waves || b->intra_rows != c->intra_waves)
        return h264g_set_error(H264G_E_STATE, "batch built for bands of %d / %d rows, the context now uses %d / %d: "
                               "create it again", b->dbk_rows, b->intra_rows, c->dbk_waves, c->intra_waves);
    for (const Batch &sb : b->sub)  // every stream before any launch
        for (const auto &sg : sb.streams)
            if (int rc = check_stream(c, sg.first, sg.second)) return rc;
    HIPCHK(hipSetDevice(c->device));
    for (int l = 0; l < 2; l++) {
        Batch &sb = b->sub[l];
        if (!sb.n_pics) continue;
        int rc = enqueue(c, batch_stages(&sb), l, sb.uses, &sb);
        if (rc) return rc;
    }
    return H264G_OK;
}

int32_t h264g_batch_destroy(h264g_ctx *c, h264g_batch *b) {
    if (!c) return h264g_set_error(H264G_E_ARG, "ctx");
    if (!b) return H264G_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    return free_batch(c, b);
}

int32_t h264g_batch_submit(h264g_ctx *c, h264g_pic *const *pics, uint32_t n) {
    h264g_batch *b = nullptr;
    int rc = h264g_batch_create(c, pics, n, &b);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    for (int l = 0; l < 2 && !rc; l++)  // used once: launched directly, not captured
        if (b->sub[l].n_pics) rc = enqueue(c, batch_stages(&b->sub[l]), l, b->sub[l].uses, nullptr);
    free_batch(c, b);                   // its blocks are reused after these launches
    return rc;
}

int32_t h264g_sync(h264g_ctx *c) {
    if (!c) return h264g_set_error(H264G_E_ARG, "ctx");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(sync_lanes(c));
    uint32_t err = 0;
    HIPCHK(hipMemcpy(&err, c->d_err, 4, hipMemcpyDeviceToHost));
    if (err) {
        HIPCHK(hipMemset(c->d_err, 0, 4));
        return h264g_set_error(H264G_E_TIMEOUT, "device error word 0x%x (wavefront wait exceeded its spin limit)", err);
    }
    return H264G_OK;
}

// A fence is one event per stream (lanes and their deblocking streams), recorded after the
// work enqueued so far on it.
struct h264g_fence {
    hipEvent_t ev[4];
};

int32_t h264g_fence_record(h264g_ctx *c, h264g_fence **out) {
    if (!c || !out) return h264g_set_error(H264G_E_ARG, "NULL argument");
    *out = nullptr;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    h264g_fence *f = new (std::nothrow) h264g_fence{{nullptr, nullptr, nullptr, nullptr}};
    if (!f) return h264g_set_error(H264G_E_NOMEM, "fence");
    for (int l = 0; l < 4; l++) {
        if (hipEventCreateWithFlags(&f->ev[l], hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(f->ev[l], any_stream(c, l)) != hipSuccess) {
            for (auto e : f->ev) if (e) (void)hipEventDestroy(e);
            delete f;
            return h264g_set_error(H264G_E_DEVICE, "fence record");
        }
    }
    *out = f;
    return H264G_OK;
}

int32_t h264g_fence_query(h264g_ctx *c, h264g_fence *f) {
    if (!c || !f) return h264g_set_error(H264G_E_ARG, "NULL argument");
    for (auto ev : f->ev) {
        hipError_t e = hipEventQuery(ev);
        if (e == hipErrorNotReady) return 0;
        if (e != hipSuccess) return h264g_set_error(H264G_E_DEVICE, "fence query: %s", hipGetErrorString(e));
    }
    return 1;
}

int32_t h264g_fence_wait(h264g_ctx *c, h264g_fence *f) {
    if (!c || !f) return h264g_set_error(H264G_E_ARG, "NULL argument");
    for (auto ev : f->ev) HIPCHK(hipEventSynchronize(ev));
    return H264G_OK;
}

int32_t h264g_fence_destroy(h264g_ctx *c, h264g_fence *f) {
    (void)c;
    if (f) {
        for (auto ev : f->ev) (void)hipEventDestroy(ev);
        delete f;
    }
    return H264G_OK;
}

int32_t h264g_set_profile(h264g_ctx *c, int32_t on) {
    if (!c) return h264g_set_error(H264G_E_ARG, "ctx");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(sync_lanes(c));
    c->profile = on != 0;
    return H264G_OK;
}

int32_t h264g_stage_times(h264g_ctx *c, float out[4]) {
    if (!c || !out || !c->profile) return h264g_set_error(H264G_E_STATE, "profiling is off");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(sync_lanes(c));
    out[0] = out[1] = out[2] = out[3] = 0.f;
    const size_t runs = c->ev_used / 6;
    for (size_t r = 0; r < runs; r++)
        for (int i = 0; i < 4; i++) {  // inter, intra, bS: consecutive events; deblocking: events 4, 5
            float ms = 0.f;
            const size_t a = 6 * r + (i < 3 ? i : 4);
            HIPCHK(hipEventElapsedTime(&ms, c->ev[a], c->ev[a + 1]));
            out[i] += ms;
        }
    c->ev_used = 0;
    return (int32_t)runs;
}

int32_t h264g_calibration_copy(h264g_ctx *c, uint64_t bytes, int32_t reps) {
    if (!c || bytes == 0 || bytes % 16 || reps < 1) return h264g_set_error(H264G_E_ARG, "calibration copy arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(hipSetDevice(c->device));
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return h264g_set_error(H264G_E_NOMEM, "calibration buffer");
    if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return h264g_set_error(H264G_E_NOMEM, "calibration buffer"); }
    int rc = 0;
    if (hipMemsetAsync(a, 1, bytes, c->stream) != hipSuccess) rc = h264g_set_error(H264G_E_DEVICE, "calibration memset");
    for (int r = 0; r < reps && !rc; r++) {
        int e = h264g_launch_copy((const uint8_t *)a, (uint8_t *)b, bytes, c->stream);
        if (e) rc = h264g_set_error(H264G_E_DEVICE, "copy launch: %s", hipGetErrorString((hipError_t)e));
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = h264g_set_error(H264G_E_DEVICE, "calibration sync");
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int32_t h264g_slot_readback(h264g_ctx *c, uint32_t sid, int32_t slot, uint8_t *y, uint8_t *cb, uint8_t *cr) {
    if (!c || !y || !cb || !cr) return h264g_set_error(H264G_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (sid >= c->streams.size() || !c->streams[sid].open) return h264g_set_error(H264G_E_ARG, "stream %u", sid);
    Stream &s = c->streams[sid];
    if (slot < 0 || slot >= s.desc.dpb_slots) return h264g_set_error(H264G_E_ARG, "slot %d", slot);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(sync_lane(c, s.lane));  // the lane (and its deblocking stream) that writes the stream
    const uint8_t *base = s.dpb + (uint64_t)slot * s.slot_bytes;
    size_t ny = (size_t)s.W * s.H * 256, nc = (size_t)s.W * s.H * 64;
    HIPCHK(hipMemcpy(y, base, ny, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cb, base + ny, nc, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cr, base + ny + nc, nc, hipMemcpyDeviceToHost));
    return H264G_OK;
}

int32_t h264g_slot_to_bgr24(h264g_ctx *c, uint32_t sid, int32_t slot, uint8_t *bgr, uint32_t stride, int32_t flip) {
    if (!c || !bgr) return h264g_set_error(H264G_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (sid >= c->streams.size() || !c->streams[sid].open) return h264g_set_error(H264G_E_ARG, "stream %u", sid);
    Stream &s = c->streams[sid];
    if (slot < 0 || slot >= s.desc.dpb_slots) return h264g_set_error(H264G_E_ARG, "slot %d", slot);
    const int wl = s.W * 16, hl = s.H * 16;
    if (stride < (uint32_t)wl * 3) return h264g_set_error(H264G_E_ARG, "stride %u < %d", stride, wl * 3);
    HIPCHK(hipSetDevice(c->device));
    const size_t bytes = (size_t)stride * hl;
    uint8_t *d = (uint8_t *)pool_get(c, bytes);
    if (!d) return h264g_set_error(H264G_E_NOMEM, "conversion buffer");
    hipStream_t ls = lane_stream(c, s.lane);  // after the stream's pictures
    HIPCHK(hipStreamSynchronize(c->dstream[s.lane]));  // and their deblocking
    int e = h264g_launch_bgr24(s.dpb + (uint64_t)slot * s.slot_bytes, wl, hl, d, stride, flip != 0, ls);
    hipError_t he = e ? (hipError_t)e : hipStreamSynchronize(ls);
    if (he == hipSuccess) he = hipMemcpy(bgr, d, bytes, hipMemcpyDeviceToHost);
    pool_put(c, d, bytes);
    if (he != hipSuccess) return h264g_set_error(H264G_E_DEVICE, "bgr24 conversion: %s", hipGetErrorString(he));
    return H264G_OK;
}

int32_t h264g_slot_upload(h264g_ctx *c, uint32_t sid, int32_t slot, const uint8_t *y, const uint8_t *cb, const uint8_t *cr) {
    if (!c || !y || !cb || !cr) return h264g_set_error(H264G_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (sid >= c->streams.size() || !c->streams[sid].open) return h264g_set_error(H264G_E_ARG, "stream %u", sid);
    Stream &s = c->streams[sid];
    if (slot < 0 || slot >= s.desc.dpb_slots) return h264g_set_error(H264G_E_ARG, "slot %d", slot);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(sync_lane(c, s.lane));
    uint8_t *base = s.dpb + (uint64_t)slot * s.slot_bytes;
    size_t ny = (size_t)s.W * s.H * 256, nc = (size_t)s.W * s.H * 64;
    HIPCHK(hipMemcpy(base, y, ny, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(base + ny, cb, nc, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(base + ny + nc, cr, nc, hipMemcpyHostToDevice));
    return H264G_OK;
}

}  // extern "C"
