// ALGO_INTERP_LORENZO parameter search (include/Falcons/api/impl/FalconsAlgoInterp.hpp:273-507).
#pragma once
#include "../../include/falcons_mi355x.h"

namespace fz {
// Chooses interp_algo / interp_direction / interp_alpha / interp_beta from test compressions of
// sampled blocks. Returns 0 when tuning is skipped (parameters unchanged), 1 when tuned, 2 when the
// reference would go on to test the Lorenzo/regression algorithm (1-D fields below ratio 50).
int tune_interp(fz_config& c, int dtype, const void* data);
// omp_get_max_threads() of the reference build decides the 2-D candidate set
// (FalconsAlgoInterp.hpp:391); FALCONS_TUNE_THREADS overrides the default of 8.
int tune_threads();
}  // namespace fz
