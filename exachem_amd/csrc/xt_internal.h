// Host-side helpers shared by the library's translation units.
#pragma once
#include <string>

namespace xt {
// Records msg as the calling thread's last error (xt_last_error) and returns code.
int set_error(int code, const std::string& msg);
}  // namespace xt
