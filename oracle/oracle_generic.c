/* CPU oracle: the generic interpolation traversal, for what the SIMD paths do not cover.
 *
 * TEST INFRASTRUCTURE ONLY (the product never links it).
 *
 * InterpolationDecomposition::compress / decompress (IDH:131-239, and the decompress loop) take this
 * traversal for 4-D fields and for integer element types (IDH:2314-2519: `is_integral_v<T>`, `N == 4`):
 *   N = 1  interpolation_1d over the line (IDH:374-419)
 *   N = 2  interpolation_1d over every 2*stride-th line along dims[0], then every stride-th line
 *          along dims[1] (IDH:2372-2386)
 *   N = 3, 4  interpolation_1d_fastest_dim_first along dims[0], dims[1], ... (IDH:436-528,
 *          2511-2519), each pass a foreach box (BlockwiseIterator.hpp:285-324)
 * with dims = the direction-th permutation of (0..N-1) in next_permutation order (IDH:330-338).
 * Every point goes through LinearQuantizer::quantize_and_overwrite / recover (LinearQuantizer.hpp:
 * 49-131): the double compare for float and double, quantize_integral / recover_integral
 * (LinearQuantizer.hpp:190-248) for integers, whose interpolators work in a wider integer type and
 * clamp back (Interpolators.hpp:33-117). The same traversal scores the parameter search's candidates
 * (TUNING::ENABLED), on each sampled block. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "falcons_oracle.h"
#include "oracle_internal.h"

#define GEN_RADIUS 32768

/* dims = the dir-th permutation of 0..N-1 in std::next_permutation order */
static int gen_perm(int N, int dir, int *dims) {
    for (int i = 0; i < N; i++) dims[i] = i;
    for (int t = 0; t < dir; t++) {
        int i = N - 2;
        while (i >= 0 && dims[i] >= dims[i + 1]) i--;
        if (i < 0) return -1;
        int j = N - 1;
        while (dims[j] <= dims[i]) j--;
        int x = dims[i]; dims[i] = dims[j]; dims[j] = x;
        for (int a = i + 1, b = N - 1; a < b; a++, b--) { x = dims[a]; dims[a] = dims[b]; dims[b] = x; }
    }
    return 0;
}

typedef struct {
    int mode;  /* 0 compress, 1 decompress */
    int err;
    double eb, ebx2, recip;
    int16_t *codes;       /* compress: written; decompress: read */
    size_t ci, ncodes;
    void *unpred;         /* compress: written; decompress: read */
    size_t ui, nunpred, unpred_cap;
    int legacy;           /* decompress of an integer stream saved with the floating uid */
} gen_q;

#define GEN_CAT2(a, b) a##_##b
#define GEN_CAT(a, b) GEN_CAT2(a, b)

/* KIND 0 floating, 1 unsigned integer, 2 signed integer. W is the interpolators' wide type. */
#define DEF_GEN(T, SFX, KIND, W, UT, TMIN, TMAX, DIGITS, FMA)                                                   \
    static inline T GEN_CAT(g_narrow, SFX)(W v) {                                                        \
        if (v < (W)(TMIN)) return (T)(TMIN);                                                             \
        if (v > (W)(TMAX)) return (T)(TMAX);                                                             \
        return (T)v;                                                                                     \
    }                                                                                                    \
    static inline T GEN_CAT(g_lin, SFX)(T a, T b) {                                                      \
        if (KIND) return GEN_CAT(g_narrow, SFX)(((W)a + (W)b) / 2);                                      \
        return (T)((a + b) / 2);                                                                         \
    }                                                                                                    \
    static inline T GEN_CAT(g_lin1, SFX)(T a, T b) {                                                     \
        if (KIND) return GEN_CAT(g_narrow, SFX)((-(W)a + 3 * (W)b) / 2);                                 \
        return (T)FMA(-(T)0.5, a, (T)1.5 * b);                                                           \
    }                                                                                                    \
    static inline T GEN_CAT(g_q1, SFX)(T a, T b, T c) {                                                  \
        if (KIND) return GEN_CAT(g_narrow, SFX)((3 * (W)a + 6 * (W)b - (W)c) / 8);                      \
        T t = (T)FMA((T)6, b, -c); t = (T)FMA((T)3, a, t); return (T)(t * (T)0.125);                    \
    }                                                                                                    \
    static inline T GEN_CAT(g_q2, SFX)(T a, T b, T c) {                                                  \
        if (KIND) return GEN_CAT(g_narrow, SFX)((-(W)a + 6 * (W)b + 3 * (W)c) / 8);                     \
        T t = (T)FMA((T)6, b, -a); t = (T)FMA((T)3, c, t); return (T)(t * (T)0.125);                    \
    }                                                                                                    \
    static inline T GEN_CAT(g_q3, SFX)(T a, T b, T c) {                                                  \
        if (KIND) return GEN_CAT(g_narrow, SFX)((3 * (W)a - 10 * (W)b + 15 * (W)c) / 8);                \
        return (T)((3 * a - 10 * b + 15 * c) / 8);                                                       \
    }                                                                                                    \
    static inline T GEN_CAT(g_cub, SFX)(T a, T b, T c, T d) {                                            \
        if (KIND) return GEN_CAT(g_narrow, SFX)((-(W)a + 9 * (W)b + 9 * (W)c - (W)d) / 16);             \
        T t = (T)((b + c) * (T)9); t = (T)(t - a); t = (T)(t - d); return (T)(t * (T)0.0625);            \
    }                                                                                                    \
    /* integral_distance (LinearQuantizer.hpp:192-198): |lhs - rhs| in the unsigned type */            \
    static inline uint64_t GEN_CAT(g_dist, SFX)(T l, T r) {                                              \
        return l < r ? (uint64_t)(UT)((UT)r - (UT)l) : (uint64_t)(UT)((UT)l - (UT)r);                    \
    }                                                                                                    \
    /* recover_integral (LinearQuantizer.hpp:200-234); 0 when the value does not fit */               \
    static inline int GEN_CAT(g_rec_int, SFX)(const gen_q *q, T pred, int qi, T *out) {                  \
        if (qi == 0) { *out = pred; return 1; }                                                          \
        const double ub = ldexp(1.0, DIGITS);                                                            \
        if (KIND == 1) {                                                                                 \
            const unsigned mag = (unsigned)(qi < 0 ? -qi : qi);                                          \
            const double corr = (double)mag * q->ebx2;                                                   \
            if (!isfinite(corr) || corr < 0) return 0;                                                   \
            const double rc = qi >= 0 ? floor(corr) : ceil(corr);                                        \
            if (rc >= ub) return 0;                                                                      \
            const T ic = (T)rc;                                                                          \
            if (qi >= 0) {                                                                               \
                if (ic > (T)((T)(TMAX) - pred)) return 0;                                                \
                *out = (T)(pred + ic);                                                                   \
            } else {                                                                                     \
                if (ic > pred) return 0;                                                                 \
                *out = (T)(pred - ic);                                                                   \
            }                                                                                            \
            return 1;                                                                                    \
        }                                                                                                \
        const double v = fma((double)qi, q->ebx2, (double)pred);                                         \
        if (!isfinite(v) || v < -ub || v >= ub) return 0;                                                \
        *out = (T)v;                                                                                     \
        return 1;                                                                                        \
    }                                                                                                    \
    /* LinearQuantizer::quantize_and_overwrite (compress) or recover (decompress) of one point */      \
    static inline void GEN_CAT(g_point, SFX)(gen_q *q, T *d, T pred) {                                   \
        if (q->mode == 0) {                                                                              \
            int code = -GEN_RADIUS;                                                                      \
            if (KIND) {                                                                                  \
                const int neg = *d < pred;                                                               \
                const uint64_t dist = GEN_CAT(g_dist, SFX)(*d, pred);                                    \
                const double r = nearbyint((double)dist * q->recip);                                     \
                if (isfinite(r) && r >= 0 && r < GEN_RADIUS) {                                           \
                    const int m = (int)r;                                                                \
                    const int qi = neg ? -m : m;                                                         \
                    T dec;                                                                               \
                    if (GEN_CAT(g_rec_int, SFX)(q, pred, qi, &dec) &&                                    \
                        (double)GEN_CAT(g_dist, SFX)(*d, dec) <= q->eb) {                                \
                        code = qi;                                                                       \
                        *d = dec;                                                                        \
                    }                                                                                    \
                }                                                                                        \
            } else {                                                                                     \
                T diff = (T)(*d - pred);                                                                 \
                double r = nearbyint((double)diff * q->recip);                                           \
                if (r > -GEN_RADIUS && r < GEN_RADIUS) {                                                 \
                    const int qi = (int)r;                                                               \
                    T dec = (T)fma((double)qi, q->ebx2, (double)pred);                                   \
                    T err = (T)(dec - *d);                                                               \
                    if ((double)err >= -q->eb && (double)err <= q->eb) {                                 \
                        code = qi;                                                                       \
                        *d = dec;                                                                        \
                    }                                                                                    \
                }                                                                                        \
            }                                                                                            \
            q->codes[q->ci++] = (int16_t)code;                                                           \
            if (code == -GEN_RADIUS) {                                                                   \
                if (q->ui < q->unpred_cap) ((T *)q->unpred)[q->ui] = *d;                                 \
                q->ui++;                                                                                 \
            }                                                                                            \
        } else {                                                                                         \
            if (q->ci >= q->ncodes) { q->err = 1; return; }                                              \
            const int code = q->codes[q->ci++];                                                          \
            if (code == -GEN_RADIUS) {                                                                   \
                if (q->ui >= q->nunpred) { q->err = 1; return; }                                         \
                *d = ((const T *)q->unpred)[q->ui++];                                                    \
            } else if (KIND && !q->legacy) {                                                             \
                if (!GEN_CAT(g_rec_int, SFX)(q, pred, code, d)) q->err = 2;                              \
            } else {                                                                                     \
                *d = (T)fma((double)code, q->ebx2, (double)pred);                                        \
            }                                                                                            \
        }                                                                                                \
    }                                                                                                    \
    /* interpolation_1d (IDH:374-419) */                                                               \
    static void GEN_CAT(g_1d, SFX)(gen_q *q, T *x, size_t begin, size_t end, size_t s, int cubic) {      \
        size_t n = (end - begin) / s + 1;                                                                \
        if (n <= 1) return;                                                                              \
        if (!cubic || n < 5) {                                                                           \
            for (size_t i = 1; i + 1 < n; i += 2) {                                                      \
                T *d = x + begin + i * s;                                                                \
                GEN_CAT(g_point, SFX)(q, d, GEN_CAT(g_lin, SFX)(d[-(long)s], d[s]));                    \
            }                                                                                            \
            if (n % 2 == 0) {                                                                            \
                T *d = x + begin + (n - 1) * s;                                                          \
                if (n < 4) GEN_CAT(g_point, SFX)(q, d, d[-(long)s]);                                     \
                else GEN_CAT(g_point, SFX)(q, d, GEN_CAT(g_lin1, SFX)(d[-3 * (long)s], d[-(long)s]));   \
            }                                                                                            \
        } else {                                                                                         \
            size_t i;                                                                                    \
            T *d;                                                                                        \
            for (i = 3; i + 3 < n; i += 2) {                                                             \
                d = x + begin + i * s;                                                                   \
                GEN_CAT(g_point, SFX)(q, d, GEN_CAT(g_cub, SFX)(d[-3 * (long)s], d[-(long)s], d[s], d[3 * s])); \
            }                                                                                            \
            d = x + begin + s;                                                                           \
            GEN_CAT(g_point, SFX)(q, d, GEN_CAT(g_q1, SFX)(d[-(long)s], d[s], d[3 * s]));               \
            d = x + begin + i * s;                                                                       \
            GEN_CAT(g_point, SFX)(q, d, GEN_CAT(g_q2, SFX)(d[-3 * (long)s], d[-(long)s], d[s]));        \
            if (n % 2 == 0) {                                                                            \
                d = x + begin + (n - 1) * s;                                                             \
                GEN_CAT(g_point, SFX)(q, d, GEN_CAT(g_q3, SFX)(d[-5 * (long)s], d[-3 * (long)s], d[-(long)s])); \
            }                                                                                            \
        }                                                                                                \
    }                                                                                                    \
    /* one foreach box of interpolation_1d_fastest_dim_first (4 slots, unused ones of extent 1) */     \
    static void GEN_CAT(g_box, SFX)(gen_q *q, T *x, const size_t *b, const size_t *e, const size_t *st,  \
                                   const size_t *off, long S, size_t n, int rule, size_t bnd) {          \
        for (size_t i0 = b[0]; i0 < e[0]; i0 += st[0])                                                   \
            for (size_t i1 = b[1]; i1 < e[1]; i1 += st[1])                                               \
                for (size_t i2 = b[2]; i2 < e[2]; i2 += st[2])                                           \
                    for (size_t i3 = b[3]; i3 < e[3]; i3 += st[3]) {                                     \
                        T *d = x + i0 * off[0] + i1 * off[1] + i2 * off[2] + i3 * off[3];                \
                        T p;                                                                             \
                        switch (rule) {                                                                  \
                            case 0: p = GEN_CAT(g_lin, SFX)(d[-S], d[S]); break;                         \
                            case 1: p = (n < 3) ? d[-S] : GEN_CAT(g_lin1, SFX)(d[-2 * S], d[-S]); break; \
                            case 2: p = GEN_CAT(g_cub, SFX)(d[-3 * S], d[-S], d[S], d[3 * S]); break;    \
                            default:                                                                     \
                                if (bnd >= 3) {                                                          \
                                    if (bnd + 3 < n) p = GEN_CAT(g_cub, SFX)(d[-3 * S], d[-S], d[S], d[3 * S]); \
                                    else if (bnd + 1 < n) p = GEN_CAT(g_q2, SFX)(d[-3 * S], d[-S], d[S]); \
                                    else p = GEN_CAT(g_lin1, SFX)(d[-3 * S], d[-S]);                     \
                                } else {                                                                 \
                                    if (bnd + 3 < n) p = GEN_CAT(g_q1, SFX)(d[-S], d[S], d[3 * S]);      \
                                    else if (bnd + 1 < n) p = GEN_CAT(g_lin, SFX)(d[-S], d[S]);          \
                                    else p = d[-S];                                                      \
                                }                                                                        \
                        }                                                                                \
                        GEN_CAT(g_point, SFX)(q, d, p);                                                  \
                    }                                                                                    \
    }                                                                                                    \
    /* interpolation_1d_fastest_dim_first (IDH:436-528) along axis pd of a 4-slot field */              \
    static void GEN_CAT(g_fdf, SFX)(gen_q *q, T *x, const size_t *D, int pd, size_t *strides, size_t st, \
                                   int cubic) {                                                          \
        size_t n = (D[pd] - 1) / st + 1;                                                                 \
        if (n <= 1) return;                                                                              \
        size_t off[4] = {D[1] * D[2] * D[3], D[2] * D[3], D[3], 1}, b[4] = {0, 0, 0, 0};                \
        size_t e[4] = {D[0], D[1], D[2], D[3]};                                                          \
        const size_t S = st * off[pd];                                                                   \
        off[pd] = S;                                                                                     \
        strides[pd] = 2;                                                                                 \
        if (!cubic) {                                                                                    \
            b[pd] = 1; e[pd] = n - 1;                                                                    \
            GEN_CAT(g_box, SFX)(q, x, b, e, strides, off, (long)S, n, 0, 0);                             \
            if (n % 2 == 0) { b[pd] = n - 1; e[pd] = n; GEN_CAT(g_box, SFX)(q, x, b, e, strides, off, (long)S, n, 1, 0); } \
        } else {                                                                                         \
            b[pd] = 3; e[pd] = (n >= 3) ? n - 3 : 0;                                                     \
            GEN_CAT(g_box, SFX)(q, x, b, e, strides, off, (long)S, n, 2, 0);                             \
            size_t bl[4]; int nb = 0;                                                                    \
            bl[nb++] = 1;                                                                                \
            if (n % 2 == 1 && n > 3) bl[nb++] = n - 2;                                                   \
            if (n % 2 == 0 && n > 4) bl[nb++] = n - 3;                                                   \
            if (n % 2 == 0 && n > 2) bl[nb++] = n - 1;                                                   \
            for (int t = 0; t < nb; t++) {                                                               \
                b[pd] = bl[t]; e[pd] = bl[t] + 1;                                                        \
                GEN_CAT(g_box, SFX)(q, x, b, e, strides, off, (long)S, n, 3, bl[t]);                     \
            }                                                                                            \
        }                                                                                                \
    }                                                                                                    \
    /* InterpolationDecomposition::compress / decompress with the generic traversal (IDH:131-239) */   \
    static int GEN_CAT(g_run, SFX)(gen_q *q, int N, const size_t *Din, int cubic, int dir, long anchor_in, \
                                   double alpha, double beta, double eb, T *x, size_t *anchor_out) {     \
        size_t anchor = anchor_in < 0 ? 0 : (size_t)anchor_in;                                           \
        int L = -1, ua = 0;                                                                              \
        for (int i = 0; i < N; i++) {                                                                    \
            const double lg = ceil(log2((double)Din[i]));                                                \
            if (L < lg) L = (int)lg;                                                                     \
            if (Din[i] > anchor) ua = 1;                                                                 \
        }                                                                                                \
        if (!ua) anchor = 0;                                                                             \
        if (anchor > 0) { int m = (int)log2((double)anchor) + 1; if (m <= L) L = m; }                    \
        if (anchor_out) *anchor_out = anchor;                                                            \
        size_t D[4] = {1, 1, 1, 1};                                                                      \
        for (int i = 0; i < N; i++) D[4 - N + i] = Din[i];                                               \
        int dims[4];                                                                                     \
        if (gen_perm(N, dir, dims)) return -1;                                                           \
        q->eb = eb; q->ebx2 = 2 * eb; q->recip = 1.0 / q->ebx2;                                          \
        if (anchor == 0) {                                                                               \
            GEN_CAT(g_point, SFX)(q, x, (T)0);                                                           \
        } else {                                                                                         \
            /* build_anchor_grid / recover_anchor_grid (IDH:341-359) */                                  \
            for (size_t a = 0; a < D[0]; a += (N > 3 ? anchor : 1))                                      \
                for (size_t b = 0; b < D[1]; b += (N > 2 ? anchor : 1))                                  \
                    for (size_t c = 0; c < D[2]; c += (N > 1 ? anchor : 1))                              \
                        for (size_t e = 0; e < D[3]; e += anchor) {                                      \
                            T *d = x + ((a * D[1] + b) * D[2] + c) * D[3] + e;                           \
                            if (q->mode == 0) {                                                          \
                                if (q->ui < q->unpred_cap) ((T *)q->unpred)[q->ui] = *d;                 \
                                q->ui++;                                                                 \
                            } else {                                                                     \
                                if (q->ui >= q->nunpred) { q->err = 1; return -2; }                      \
                                *d = ((const T *)q->unpred)[q->ui++];                                    \
                            }                                                                            \
                        }                                                                                \
            L--;                                                                                         \
        }                                                                                                \
        for (int level = L; level > 0 && level <= L; level--) {                                          \
            double cur = eb;                                                                             \
            if (alpha < 0) cur = (level >= 3) ? eb * 0.5 : eb;                                           \
            else if (alpha >= 1) { double r = pow(alpha, level - 1); if (r > beta) r = beta; cur = eb / r; } \
            q->eb = cur; q->ebx2 = 2 * cur; q->recip = 1.0 / q->ebx2;                                    \
            const size_t st = (size_t)1 << (level - 1);                                                  \
            if (N == 1) {                                                                                \
                GEN_CAT(g_1d, SFX)(q, x, 0, Din[0] - 1, st, cubic);                                      \
            } else if (N == 2) {                                                                         \
                const size_t off[2] = {Din[1], 1};                                                       \
                const int d0 = dims[0], d1 = dims[1];                                                    \
                for (size_t j = 0; j <= Din[d1] - 1; j += 2 * st) {                                      \
                    const size_t bo = j * off[d1];                                                       \
                    GEN_CAT(g_1d, SFX)(q, x, bo, bo + (Din[d0] - 1) * off[d0], st * off[d0], cubic);     \
                }                                                                                        \
                for (size_t i = 0; i <= Din[d0] - 1; i += st) {                                          \
                    const size_t bo = i * off[d0];                                                       \
                    GEN_CAT(g_1d, SFX)(q, x, bo, bo + (Din[d1] - 1) * off[d1], st * off[d1], cubic);     \
                }                                                                                        \
            } else {                                                                                     \
                size_t strides[4] = {1, 1, 1, 1};                                                        \
                const int o = 4 - N; /* field axis a sits in slot o + a */                               \
                strides[o + dims[0]] = 1;                                                                \
                for (int i = 1; i < N; i++) strides[o + dims[i]] = 2 * st;                               \
                GEN_CAT(g_fdf, SFX)(q, x, D, o + dims[0], strides, st, cubic);                           \
                for (int i = 1; i < N; i++) {                                                            \
                    strides[o + dims[i - 1]] = st;                                                       \
                    GEN_CAT(g_fdf, SFX)(q, x, D, o + dims[i], strides, st, cubic);                       \
                }                                                                                        \
            }                                                                                            \
            if (q->err) return -2;                                                                       \
        }                                                                                                \
        return q->err ? -2 : 0;                                                                          \
    }

DEF_GEN(float, f32, 0, double, uint32_t, -1, 1, 24, fmaf)
DEF_GEN(double, f64, 0, double, uint64_t, -1, 1, 53, fma)
DEF_GEN(uint8_t, u8, 1, int32_t, uint8_t, 0, UINT8_MAX, 8, fma)
DEF_GEN(int8_t, i8, 2, int32_t, uint8_t, INT8_MIN, INT8_MAX, 7, fma)
DEF_GEN(uint16_t, u16, 1, int32_t, uint16_t, 0, UINT16_MAX, 16, fma)
DEF_GEN(int16_t, i16, 2, int32_t, uint16_t, INT16_MIN, INT16_MAX, 15, fma)
DEF_GEN(uint32_t, u32, 1, int64_t, uint32_t, 0, UINT32_MAX, 32, fma)
DEF_GEN(int32_t, i32, 2, int64_t, uint32_t, INT32_MIN, INT32_MAX, 31, fma)
DEF_GEN(uint64_t, u64, 1, __int128, uint64_t, 0, UINT64_MAX, 64, fma)
DEF_GEN(int64_t, i64, 2, __int128, uint64_t, INT64_MIN, INT64_MAX, 63, fma)

size_t fo_elem_size(int dtype) {
    static const size_t es[10] = {4, 8, 1, 1, 2, 2, 4, 4, 8, 8};
    return (dtype >= 0 && dtype < 10) ? es[dtype] : 0;
}

static int gen_dispatch(gen_q *q, int dtype, int N, const size_t *dims, int cubic, int dir, long anchor,
                        double alpha, double beta, double eb, void *x, size_t *anchor_out) {
    switch (dtype) {
        case 0: return g_run_f32(q, N, dims, cubic, dir, anchor, alpha, beta, eb, (float *)x, anchor_out);
        case 1: return g_run_f64(q, N, dims, cubic, dir, anchor, alpha, beta, eb, (double *)x, anchor_out);
        case 2: return g_run_u8(q, N, dims, cubic, dir, anchor, alpha, beta, eb, (uint8_t *)x, anchor_out);
        case 3: return g_run_i8(q, N, dims, cubic, dir, anchor, alpha, beta, eb, (int8_t *)x, anchor_out);
        case 4: return g_run_u16(q, N, dims, cubic, dir, anchor, alpha, beta, eb, (uint16_t *)x, anchor_out);
        case 5: return g_run_i16(q, N, dims, cubic, dir, anchor, alpha, beta, eb, (int16_t *)x, anchor_out);
        case 6: return g_run_u32(q, N, dims, cubic, dir, anchor, alpha, beta, eb, (uint32_t *)x, anchor_out);
        case 7: return g_run_i32(q, N, dims, cubic, dir, anchor, alpha, beta, eb, (int32_t *)x, anchor_out);
        case 8: return g_run_u64(q, N, dims, cubic, dir, anchor, alpha, beta, eb, (uint64_t *)x, anchor_out);
        case 9: return g_run_i64(q, N, dims, cubic, dir, anchor, alpha, beta, eb, (int64_t *)x, anchor_out);
        default: return -1;
    }
}

int fo_gen_compress(int dtype, int N, const size_t *dims, int cubic, int dir, long anchor, double alpha,
                    double beta, double eb, void *data, int16_t *codes, size_t *ncodes, void *unpred,
                    size_t unpred_cap, size_t *nunpred, size_t *anchor_out) {
    gen_q q;
    memset(&q, 0, sizeof q);
    q.mode = 0;
    q.codes = codes;
    q.unpred = unpred;
    q.unpred_cap = unpred_cap;
    const int rc = gen_dispatch(&q, dtype, N, dims, cubic, dir, anchor, alpha, beta, eb, data, anchor_out);
    *ncodes = q.ci;
    *nunpred = q.ui;
    return rc;
}

int fo_gen_decompress(int dtype, int N, const size_t *dims, int cubic, int dir, long anchor, double alpha,
                      double beta, double eb, int legacy, const int16_t *codes, size_t ncodes, const void *unpred,
                      size_t nunpred, void *out) {
    gen_q q;
    memset(&q, 0, sizeof q);
    q.mode = 1;
    q.codes = (int16_t *)codes;
    q.ncodes = ncodes;
    q.unpred = (void *)unpred;
    q.nunpred = nunpred;
    q.legacy = legacy;
    const int rc = gen_dispatch(&q, dtype, N, dims, cubic, dir, anchor, alpha, beta, eb, out, NULL);
    if (rc == 0 && (q.ci != ncodes || q.ui != nunpred)) return -3;
    return q.err == 2 ? -4 : rc;
}

/* LinearQuantizer on one integer point (for the pinning tests): returns the code; *out is the
 * overwritten value; *escaped = 1 when the point went to the unpredictable list. */
#define DEF_QPOINT(T, SFX)                                                                              \
    int GEN_CAT(fo_quantize_point, SFX)(T ori, T pred, double eb, T *out, int *escaped) {              \
        gen_q q;                                                                                         \
        int16_t code;                                                                                    \
        T un;                                                                                            \
        memset(&q, 0, sizeof q);                                                                         \
        q.eb = eb; q.ebx2 = 2 * eb; q.recip = 1.0 / q.ebx2;                                              \
        q.codes = &code; q.unpred = &un; q.unpred_cap = 1;                                               \
        T d = ori;                                                                                       \
        GEN_CAT(g_point, SFX)(&q, &d, pred);                                                             \
        *out = d;                                                                                        \
        *escaped = q.ui == 1;                                                                            \
        return code;                                                                                     \
    }                                                                                                    \
    int GEN_CAT(fo_recover_point, SFX)(T pred, int code, double eb, T *out) {                            \
        gen_q q;                                                                                         \
        memset(&q, 0, sizeof q);                                                                         \
        q.eb = eb; q.ebx2 = 2 * eb; q.recip = 1.0 / q.ebx2;                                              \
        return GEN_CAT(g_rec_int, SFX)(&q, pred, code, out);                                             \
    }                                                                                                    \
    T GEN_CAT(fo_interp_point, SFX)(int kind, T a, T b, T c, T d) {                                      \
        switch (kind) {                                                                                  \
            case 0: return GEN_CAT(g_lin, SFX)(a, b);                                                    \
            case 1: return GEN_CAT(g_lin1, SFX)(a, b);                                                   \
            case 2: return GEN_CAT(g_q1, SFX)(a, b, c);                                                  \
            case 3: return GEN_CAT(g_q2, SFX)(a, b, c);                                                  \
            case 5: return GEN_CAT(g_q3, SFX)(a, b, c);                                                  \
            default: return GEN_CAT(g_cub, SFX)(a, b, c, d);                                             \
        }                                                                                                \
    }
DEF_QPOINT(uint8_t, u8)
DEF_QPOINT(int8_t, i8)
DEF_QPOINT(uint16_t, u16)
DEF_QPOINT(int16_t, i16)
DEF_QPOINT(uint32_t, u32)
DEF_QPOINT(int32_t, i32)
DEF_QPOINT(uint64_t, u64)
DEF_QPOINT(int64_t, i64)
