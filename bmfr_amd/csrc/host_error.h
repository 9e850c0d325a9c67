// Status-code errors of the C ABI: the message of the last failing call on
// this thread (bmfr_last_error).
#pragma once
#include <cstdint>

namespace bmfr_host {
// Records the printf-style message and returns `code`.
int32_t fail(int32_t code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const char *last_error();
}  // namespace bmfr_host
