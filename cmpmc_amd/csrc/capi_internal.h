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

// Film-entry probabilities of the batched form for one AMM surface: t_L and t_T at
// |cos theta| = i / n, i = 0..n, into out[2 * (n + 1)] (L row first).
void amm_entry_table(const cmpmc_medium* substrate, const cmpmc_medium* film, int n, float* out);

}  // namespace cmpmc_internal
