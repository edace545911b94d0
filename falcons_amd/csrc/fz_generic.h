// The generic interpolation traversal (4-D fields, integer element types) and the pieces of the
// pipeline that depend on the element type beyond float and double (fz_generic.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "fz_common.h"

namespace fz {

// Bytes of one element of the Falcons_FLOAT .. Falcons_INT64 types (Config.hpp:29-38), 0 if unknown.
size_t elem_size(int dtype);

// One foreach box of interpolation_1d_fastest_dim_first over four slots (slot 0 outermost): points
// b + i*step per slot (cnt of them) at element offset sum (b + i*step)*off; rule picks the prediction
// (0 linear, 1 linear end row, 2 cubic bulk, 3 cubic boundary row bnd) along the pass axis, whose
// grid points are S elements apart, n of them.
struct GBox {
    uint64_t b[4], cnt[4], step[4], off[4];
    int64_t S;
    uint64_t n, bnd, code_base, total;
    int32_t rule, pad;
};

// interpolation_1d over a family of lines: line l starts at first + l*line_stride, n grid points S
// apart, n/2 targets each.
struct GLines {
    uint64_t first, line_stride, nlines;
    int64_t S;
    uint64_t n, code_base;
    int32_t cubic, pad;
};

struct GStep {
    int kind;  // 0 first point, 1 box, 2 lines
    LevelEb L;
    GBox box;
    GLines lines;
};

// The anchor grid: cnt points per slot, str elements apart.
struct GAnchors {
    uint64_t cnt[4] = {1, 1, 1, 1}, str[4] = {0, 0, 0, 0};
    uint64_t total = 0;
};

// Decompression side: sorted stream positions of the escaped codes, their values in that order,
// whether an integer stream uses the floating reconstruction (saved with the floating uid), and a
// device flag raised when an integer reconstruction does not fit its type.
struct GenEsc {
    const uint64_t* idx = nullptr;
    uint64_t n = 0;
    const void* vals = nullptr;
    int legacy = 0;
    int* err = nullptr;
};

struct GenPlan {
    int N = 0;
    uint64_t anchor = 0;
    int levels = 0;
    uint64_t ncodes = 0;
    GAnchors anchors;
    std::vector<GStep> steps;
};

LevelEb gen_level_eb(double eb);
// The traversal of a field (or of one block) of N <= 4 dims as launch steps.
GenPlan gen_plan(int N, const uint64_t* dims, long anchor, int cubic, int dir, double alpha, double beta, double eb);
// Runs the steps on nblk blocks of per elements (codes ncb apart); compress writes codes and the
// escaped originals (escv, indexed like the codes), decompress reads the codes and e.
void gen_traverse(int dtype, bool decomp, const std::vector<GStep>& steps, uint64_t nblk, uint64_t per, uint64_t ncb,
                  void* data, int16_t* codes, void* escv, const GenEsc& e, hipStream_t st);
void gen_anchors(int dtype, bool decomp, void* data, void* unpred, const GAnchors& A, hipStream_t st);
// REL / PSNR / combined bounds of an integer field (data_range on the element type).
size_t gen_minmax_scratch_bytes();
void gen_abs_error_bound(int dtype, const void* data, uint64_t n, int mode, double abs_in, double rel,
                         double psnr_factor, void* scratch, double* eb_out, hipStream_t st);
// Escaped originals in code order, for any element size.
void gen_gather_escapes(int dtype, const uint64_t* idx, const uint64_t* nsel_dev, uint64_t max_n, const void* escval,
                        void* out, hipStream_t st);

}  // namespace fz
