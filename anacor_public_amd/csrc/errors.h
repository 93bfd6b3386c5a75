// Library-wide error state behind anacor_last_error(): the message is kept
// per calling thread, the code is the last one set by any thread.
#pragma once

namespace anacor {
void set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
}  // namespace anacor
