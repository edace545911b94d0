/* CPU oracle: interpolation decomposition + linear quantizer (test infrastructure only).
 * The loop nests live in oracle_interp_impl.h, instantiated for float and double. */
#include <math.h>
#include <stdlib.h>

#include "falcons_oracle.h"
#include "oracle_internal.h"

/* 1: AVX2 build semantics (f32 lane compares), 0: scalar build (every point through
 * LinearQuantizer's double compare, Interpolation_quantizer.inl:2620-2848) */
static int fo_g_avx2 = 1;
void fo_set_simd(int avx2) { fo_g_avx2 = avx2 ? 1 : 0; }
int fo_get_simd(void) { return fo_g_avx2; }

static inline float fo_fma_f32(float a, float b, float c) { return fmaf(a, b, c); }
static inline double fo_fma_f64(double a, double b, double c) { return fma(a, b, c); }

#define FT float
#define SFX f32
#include "oracle_interp_impl.h"
#undef FT
#undef SFX
#define FT double
#define SFX f64
#include "oracle_interp_impl.h"
#undef FT
#undef SFX

int fo_interp_plan(const size_t *dims, int N, long anchor, size_t *anchor_out, int *levels) {
    size_t a = anchor < 0 ? 0 : (size_t)anchor;
    plan_f32(dims, N, &a, levels);
    *anchor_out = a;
    return 0;
}

int fo_interp_compress(int dtype, int N, const size_t *dims, int cubic, int dir, long anchor, double alpha,
                       double beta, double eb, void *data, int16_t *codes, size_t *ncodes, void *unpred,
                       size_t unpred_cap, size_t *nunpred, size_t *anchor_out) {
    int rc;
    size_t a = anchor < 0 ? 0 : (size_t)anchor;
    if (dtype == 0) {
        fo_ctx_f32 c = {0};
        c.data = (float *)data; c.codes = codes; c.unpred = (float *)unpred; c.un_cap = unpred_cap;
        rc = run_f32(&c, dims, N, cubic, dir, a, alpha, beta, eb, anchor_out);
        *ncodes = c.nq; *nunpred = c.nun;
    } else {
        fo_ctx_f64 c = {0};
        c.data = (double *)data; c.codes = codes; c.unpred = (double *)unpred; c.un_cap = unpred_cap;
        rc = run_f64(&c, dims, N, cubic, dir, a, alpha, beta, eb, anchor_out);
        *ncodes = c.nq; *nunpred = c.nun;
    }
    return rc;
}

int fo_interp_decompress(int dtype, int N, const size_t *dims, int cubic, int dir, long anchor, double alpha,
                         double beta, double eb, const int16_t *codes, size_t ncodes, const void *unpred,
                         size_t nunpred, void *out) {
    int rc;
    size_t a = anchor < 0 ? 0 : (size_t)anchor, ao;
    (void)ncodes; (void)nunpred;
    if (dtype == 0) {
        fo_ctx_f32 c = {0};
        c.data = (float *)out; c.codes = (int16_t *)codes; c.decomp = 1; c.unpred_in = (const float *)unpred;
        rc = run_f32(&c, dims, N, cubic, dir, a, alpha, beta, eb, &ao);
    } else {
        fo_ctx_f64 c = {0};
        c.data = (double *)out; c.codes = (int16_t *)codes; c.decomp = 1; c.unpred_in = (const double *)unpred;
        rc = run_f64(&c, dims, N, cubic, dir, a, alpha, beta, eb, &ao);
    }
    return rc;
}

/* Scalar helpers exposed for the known-answer tests */
float fo_interp_scalar_f32(int kind, float a, float b, float c, float d) {
    switch (kind) {
        case 0: return i_lin_f32(a, b);
        case 1: return i_lin1_f32(a, b);
        case 2: return i_q1_f32(a, b, c);
        case 3: return i_q2_f32(a, b, c);
        default: return i_cub_f32(a, b, c, d);
    }
}
double fo_interp_scalar_f64(int kind, double a, double b, double c, double d) {
    switch (kind) {
        case 0: return i_lin_f64(a, b);
        case 1: return i_lin1_f64(a, b);
        case 2: return i_q1_f64(a, b, c);
        case 3: return i_q2_f64(a, b, c);
        default: return i_cub_f64(a, b, c, d);
    }
}
/* One point through the scalar (vclass 0) or AVX2-lane (vclass 1) quantizer */
int fo_quantize_point_f32(float ori, float pred, double eb, int vclass, float *out, int *escaped) {
    float d = ori, un = 0;
    int16_t code;
    fo_ctx_f32 c = {0};
    c.data = &d; c.codes = &code; c.unpred = &un; c.un_cap = 1;
    c.eb = eb; c.ebx2 = 2 * eb; c.recip = 1.0 / c.ebx2; c.ebf = (float)eb;
    point_f32(&c, 0, pred, vclass, 0);
    *out = d;
    *escaped = (int)c.nun;
    return code;
}
int fo_quantize_point_f64(double ori, double pred, double eb, int vclass, double *out, int *escaped) {
    double d = ori, un = 0;
    int16_t code;
    fo_ctx_f64 c = {0};
    c.data = &d; c.codes = &code; c.unpred = &un; c.un_cap = 1;
    c.eb = eb; c.ebx2 = 2 * eb; c.recip = 1.0 / c.ebx2; c.ebf = (float)eb;
    point_f64(&c, 0, pred, vclass, 0);
    *out = d;
    *escaped = (int)c.nun;
    return code;
}
