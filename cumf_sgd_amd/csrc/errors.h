// Thread-local last-error text shared by the C-ABI translation units.
#pragma once

namespace cumf {

// Records the formatted message as the calling thread's last error and
// returns `code` (a CUMF_ERR_* value).
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

}  // namespace cumf
