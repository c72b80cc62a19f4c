// The message mpd_last_error() returns: the last failure on the calling thread,
// recorded by every translation unit of the library through set_error.
#pragma once

namespace mpd {

// Record the message for this thread and return `code`.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

}  // namespace mpd
