// status.hpp -- error reporting shared by the C ABI translation units.
#pragma once

#include <string>

namespace e2d {

// Records `msg` as the calling thread's last error (returned by
// e2d_last_error()) and returns `code`.
int report(int code, const std::string &msg);

}  // namespace e2d
