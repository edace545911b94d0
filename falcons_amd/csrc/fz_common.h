// Shared definitions of the MI355X Falcons path: pass descriptors, per-level quantizer constants,
// error handling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

#define FZ_HIP(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw fz::GpuError(std::string(#call) + " failed: " + hipGetErrorString(e_));          \
    } while (0)

namespace fz {

struct GpuError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Quantizer constants of one level (LinearQuantizer.hpp:39-43): the level's bound, twice the bound,
// its reciprocal, and the bound rounded to float for the AVX2 lane compare
// (InterpolationDecomposition.hpp:2301-2311).
struct LevelEb {
    double eb, ebx2, recip;
    float ebf;
    float pad;
};

// One interpolation pass over the whole field.
//  ROWS: a pass along a slow axis (3D z and y passes, 2D y pass). Targets are the odd grid indices g
//        of the pass axis (n points, sp elements apart); each target row runs along the fastest axis
//        (len points, sk apart); rows are grouped by an outer index o (no values, so apart).
//  LINES: a pass along the fastest axis (3D x pass, 2D x pass, 1D). Lines are enumerated by (a, b)
//        (n1 x n2, element spacings s1, s2); targets are the odd grid indices of each line.
struct PassDesc {
    int32_t type;   // 0 ROWS, 1 LINES
    int32_t cubic;  // 0 linear, 1 cubic
    int32_t skip;   // no write-back (the last pass of level 1)
    int32_t is2d;   // ROWS cubic: even-n end row coded by the vector kernel (2D) or the scalar one (3D)
    int32_t level;  // index into the LevelEb table
    int32_t pad;
    uint64_t n, sp;
    uint64_t no, so, len, sk;  // ROWS
    uint64_t n1, s1, n2, s2;   // LINES
    uint64_t code_base;        // stream index of the pass's first code
    uint64_t ncodes;           // codes this pass emits
};

enum Dtype { kF32 = 0, kF64 = 1 };

}  // namespace fz
