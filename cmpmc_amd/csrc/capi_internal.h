// Error bookkeeping shared by the C ABI translation units.
#pragma once

#include <cstdarg>
#include <cstdio>

namespace cmpmc_internal {

// Records `fmt` as the calling thread's last error and returns `code`.
int fail(int code, const char* fmt, ...);
void clear_error();

}  // namespace cmpmc_internal
