// errors.hpp — the last-error message of the C ABI (astraea_last_error), for host-side sources.
#pragma once

namespace astraea_hip {
// Records the message for astraea_last_error on the calling thread and returns `code`.
int reportError(int code, const char* fmt, ...);
}  // namespace astraea_hip
