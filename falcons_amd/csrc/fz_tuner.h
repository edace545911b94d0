// ALGO_INTERP_LORENZO parameter search (include/Falcons/api/impl/FalconsAlgoInterp.hpp:273-507).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/falcons_mi355x.h"
#include "fz_kernels.h"

namespace fz {

// Ratios of the candidates the last search evaluated, in evaluation order.
struct TuneReport {
    int n = 0;
    double ratio[16];
};

// Chooses interp_algo / interp_direction / interp_alpha / interp_beta from test compressions of
// sampled blocks of the device field d_data (c.abs_eb already absolute). Returns 0 when tuning is
// skipped (parameters unchanged), 1 when tuned, 2 when the reference would go on to test the
// Lorenzo/regression algorithm (1-D fields below ratio 50). Workspace comes from ws and is given
// back before returning.
int tune_interp(fz_config& c, int dtype, const void* d_data, DevArena& ws, hipStream_t st, TuneReport* rep);
// omp_get_max_threads() of the reference build decides the 2-D candidate set
// (FalconsAlgoInterp.hpp:391); FALCONS_TUNE_THREADS overrides the default of 8.
int tune_threads();

}  // namespace fz
