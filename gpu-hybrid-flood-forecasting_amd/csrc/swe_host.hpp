// Host-side helpers shared by the translation units of libswe_hip.
#pragma once

#include <string>

namespace swe {
// Record `msg` as the calling thread's last error (swe_last_error) and return `code`.
int set_error(int code, const std::string& msg);
}  // namespace swe
