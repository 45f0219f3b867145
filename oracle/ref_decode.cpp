// TEST INFRASTRUCTURE ONLY: nothing on the product path links or runs this file.
//
// The reference decoder, whole (CH264VideoDecoder::open on an Annex B file, every
// translation unit of the reference except its main.cpp, compiled unmodified by
// oracle/Makefile), with an output callback (the shape of main.cpp:14-47) that prints,
// per output picture in the reference's output order, one JSON line: output index,
// PicOrderCnt, slice type, decoding-order count, and the SHA-256 of the Y, Cb and Cr
// planes (stride = width, as CH264PictureBase stores them, H264PictureBase.cpp:167-179),
// prefixed with "REFPIC " (the reference prints its own diagnostics on stdout too).
//
//   ref_decode FILE [MAX_PICTURES]
//
// MAX_PICTURES stops after that many output pictures (the callback returns -1, which the
// reference treats as the end of decoding, H264VideoDecoder.cpp:383-436).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "H264VideoDecoder.h"

namespace {

// SHA-256 (FIPS 180-4), small and self-contained, so the printed digests compare
// directly with hashlib.sha256 in the tests.
struct Sha256 {
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    uint8_t buf[64];
    uint64_t len = 0;
    size_t used = 0;
    static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
    void block(const uint8_t *p) {
        static const uint32_t k[64] = {
            0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
            0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
            0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
            0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
            0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
            0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
            0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
            0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
        uint32_t w[64];
        for (int i = 0; i < 16; i++) w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 | (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
        for (int i = 16; i < 64; i++) {
            const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
        for (int i = 0; i < 64; i++) {
            const uint32_t t1 = hh + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + k[i] + w[i];
            const uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
    }
    void add(const uint8_t *p, size_t n) {
        len += n;
        while (n) {
            const size_t t = std::min(n, (size_t)64 - used);
            memcpy(buf + used, p, t);
            used += t; p += t; n -= t;
            if (used == 64) { block(buf); used = 0; }
        }
    }
    std::string hex() {
        const uint64_t bits = len * 8;
        const uint8_t one = 0x80, zero = 0;
        add(&one, 1);
        while (used != 56) add(&zero, 1);
        uint8_t l[8];
        for (int i = 0; i < 8; i++) l[i] = (uint8_t)(bits >> (56 - 8 * i));
        add(l, 8);
        char s[65];
        for (int i = 0; i < 8; i++) snprintf(s + 8 * i, 9, "%08x", h[i]);
        return s;
    }
};

std::string sha(const uint8_t *p, size_t n) {
    Sha256 s;
    s.add(p, n);
    return s.hex();
}

struct Out {
    int count = 0, limit = 0;
};

int on_picture(CH264Picture *pic, void *user, int /*error*/) {
    Out *o = (Out *)user;
    if (!pic) return -1;
    const CH264PictureBase &f = pic->m_picture_frame;
    const size_t ny = (size_t)f.PicWidthInSamplesL * f.PicHeightInSamplesL;
    const size_t nc = (size_t)f.PicWidthInSamplesC * f.PicHeightInSamplesC;
    printf("REFPIC {\"i\": %d, \"poc\": %d, \"slice_type\": %d, \"decode_index\": %d, \"width\": %d, \"height\": %d, "
           "\"y\": \"%s\", \"cb\": \"%s\", \"cr\": \"%s\"}\n",
           o->count, f.PicOrderCnt, (int)(f.m_h264_slice_header.slice_type % 5), f.m_PicNumCnt, f.PicWidthInSamplesL,
           f.PicHeightInSamplesL, sha(f.m_pic_buff_luma, ny).c_str(), sha(f.m_pic_buff_cb, nc).c_str(),
           sha(f.m_pic_buff_cr, nc).c_str());
    fflush(stdout);
    o->count++;
    return (o->limit > 0 && o->count >= o->limit) ? -1 : 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s FILE [MAX_PICTURES]\n", argv[0]);
        return 2;
    }
    Out o;
    o.limit = argc > 2 ? atoi(argv[2]) : 0;
    CH264VideoDecoder vd;
    vd.set_output_frame_callback_functuin(on_picture, &o);
    const int ret = vd.open(argv[1]);
    fprintf(stderr, "ref_decode: open returned %d after %d output pictures\n", ret, o.count);
    return 0;
}
