// TEST INFRASTRUCTURE ONLY: nothing on the product path links or runs this file.
//
// Prints the CABAC context initialisation values (m, n) the reference uses for every
// ctxIdx 0..459 (Tables 9-12 to 9-24 of the standard: 4:2:0 and 4:2:2 frame and field
// coding), from its own CH264Cabac::get_m_and_n (H264Cabac.cpp:37-1046), one line per
// ctxIdx: ctxIdx, then m n for I slices and for cabac_init_idc 0, 1, 2 ("x x" where the
// context does not exist for that slice type).  oracle/make_cabac_init.py turns it into
// the data file h264_video_decoder_demo_amd/csrc/cabac_init.inc.
#include <cstdio>

#include "H264Cabac.h"

int main() {
    CH264Cabac c;
    for (int ctx = 0; ctx <= 459; ctx++) {
        printf("%d", ctx);
        for (int k = 0; k < 4; k++) {
            int32_t m = 0, n = 0;
            const int rc = k == 0 ? c.get_m_and_n(ctx, H264_SLIECE_TYPE_I, 0, m, n)
                                  : c.get_m_and_n(ctx, H264_SLIECE_TYPE_P, k - 1, m, n);
            if (rc != 0) printf(" x x");
            else printf(" %d %d", m, n);
        }
        printf("\n");
    }
    return 0;
}
