// Error bookkeeping shared by the C ABI translation units.
#pragma once

#include <cstdarg>
#include <cstdio>

#include "cmpmc_amd.h"

namespace cmpmc_internal {

// Records `fmt` as the calling thread's last error and returns `code`.
int fail(int code, const char* fmt, ...);
void clear_error();

// C_ijlm of the crystal's cubic Voigt matrix, row-major [3][3][3][3].
void stiffness_tensor(const cmpmc_crystal& c, double out[81]);

}  // namespace cmpmc_internal
