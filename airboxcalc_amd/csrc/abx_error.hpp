// Error state of the library (abx_last_error), without the HIP headers, so
// host-only sources build with a plain host compiler.
#pragma once

#include <string>

namespace abx {

// Thread-local message of the last failure (abx_last_error).
void set_error(const std::string& msg);
void set_errorf(const char* fmt, ...);
const char* last_error();

}  // namespace abx
