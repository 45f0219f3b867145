// gen_err.cpp -- TEST INFRASTRUCTURE ONLY: error sink for the generator when it is
// linked into oracle/_ref/ref_harness without the rest of libh264g.
#include <cstdarg>
#include <cstdio>
#include <cstdint>

extern "C" int32_t h264g_set_error(int32_t code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fputc('\n', stderr);
    return code;
}
