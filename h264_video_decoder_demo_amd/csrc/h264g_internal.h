// h264g_internal.h -- shared host-side helpers of libh264g (not part of the ABI).
#pragma once
#include <stdint.h>

// Records an error message for h264g_last_error() and returns `code`.
extern "C" int32_t h264g_set_error(int32_t code, const char *fmt, ...);
