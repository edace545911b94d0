// Predictors and the linear quantizer of the interpolation path, shared by the compression kernels
// and the tuner's kernels. Compiled with -ffp-contract=off: every multiply/add rounds on its own
// unless written as fma, as in the reference.
#pragma once
#include <hip/hip_runtime.h>

#include "fz_common.h"

namespace fz {

// ---- predictors (Interpolators.hpp:48-117 and the vector forms in Interpolation_quantizer.inl) ----
template <typename T>
__device__ __forceinline__ T p_lin(T a, T b) { return (a + b) * (T)0.5; }
template <typename T>
__device__ __forceinline__ T p_cub(T a, T b, T c, T d) { return (((b + c) * (T)9) - a - d) * (T)0.0625; }
template <typename T>
__device__ __forceinline__ T p_l1(T a, T b) { return fma(-(T)0.5, a, b * (T)1.5); }
template <typename T>
__device__ __forceinline__ T p_q1(T a, T b, T c) { return fma((T)3, a, fma((T)6, b, -c)) * (T)0.125; }
template <typename T>
__device__ __forceinline__ T p_q2(T a, T b, T c) { return fma((T)3, c, fma((T)6, b, -a)) * (T)0.125; }

// Quantize one point. vlane: the point is coded by an AVX2 lane in the reference, whose error check
// compares the float error with the float bound (Interpolation_quantizer.inl:1200-1239); otherwise
// the double compare of LinearQuantizer::quantize_and_overwrite (LinearQuantizer.hpp:49-82).
template <typename T>
__device__ __forceinline__ int quantize(T ori, T pred, const LevelEb& L, bool vlane, T& dec) {
    const T diff = ori - pred;
    const double q = rint((double)diff * L.recip);
    if (!(q > -32768.0 && q < 32768.0)) return -32768;
    dec = (T)fma(q, L.ebx2, (double)pred);
    const T err = dec - ori;
    bool ok;
    if (sizeof(T) == 4 && vlane) {
        const float e = (float)err;
        ok = (e >= -L.ebf) && (e <= L.ebf);
    } else {
        ok = ((double)err >= -L.eb) && ((double)err <= L.eb);
    }
    return ok ? (int)q : -32768;
}

}  // namespace fz
