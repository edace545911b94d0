/* CPU oracle: ALGO_INTERP_LORENZO parameter search (api/impl/FalconsAlgoInterp.hpp:100-226,
 * 273-507; utils/Sample.hpp; the TUNING::ENABLED traversal of InterpolationDecomposition.hpp:
 * interpolation_1d (:374-419) and interpolation_1d_fastest_dim_first (:436-528)).
 * Test infrastructure only. */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle_zstd.h"

#include "falcons_oracle.h"
#include "oracle_internal.h"

#define TT_CAT2(a, b) a##_##b
#define TT_CAT(a, b) TT_CAT2(a, b)

/* The tuning traversal quantizes every point with LinearQuantizer::quantize_and_overwrite. */
#define DEF_TUNE(FT, SFX, FMA)                                                                                \
    typedef struct { FT *d; int16_t *codes; size_t nq; FT *un; size_t nun; double eb, ebx2, recip; } TT_CAT(tq, SFX); \
    static void TT_CAT(tq_point, SFX)(TT_CAT(tq, SFX) *q, FT *d, FT pred) {                                  \
        FT diff = *d - pred;                                                                                 \
        double r = nearbyint((double)diff * q->recip);                                                       \
        int ok = (r > -32768.0 && r < 32768.0);                                                              \
        FT dec = 0;                                                                                          \
        if (ok) {                                                                                            \
            dec = (FT)fma(r, q->ebx2, (double)pred);                                                         \
            FT err = dec - *d;                                                                               \
            ok = ((double)err >= -q->eb && (double)err <= q->eb);                                            \
        }                                                                                                    \
        if (ok) { q->codes[q->nq++] = (int16_t)(int)r; *d = dec; }                                           \
        else { q->codes[q->nq++] = -32768; q->un[q->nun++] = *d; }                                           \
    }                                                                                                        \
    static FT TT_CAT(t_lin, SFX)(FT a, FT b) { return (a + b) / 2; }                                         \
    static FT TT_CAT(t_lin1, SFX)(FT a, FT b) { return FMA(-(FT)0.5, a, (FT)1.5 * b); }                      \
    static FT TT_CAT(t_q1, SFX)(FT a, FT b, FT c) { FT t = FMA((FT)6, b, -c); t = FMA((FT)3, a, t); return t * (FT)0.125; } \
    static FT TT_CAT(t_q2, SFX)(FT a, FT b, FT c) { FT t = FMA((FT)6, b, -a); t = FMA((FT)3, c, t); return t * (FT)0.125; } \
    static FT TT_CAT(t_q3, SFX)(FT a, FT b, FT c) { return (3 * a - 10 * b + 15 * c) / 8; }                  \
    static FT TT_CAT(t_cub, SFX)(FT a, FT b, FT c, FT d) { FT t = (b + c) * (FT)9; t = t - a; t = t - d; return t * (FT)0.0625; } \
    /* interpolation_1d (IDH:374-419) */                                                                     \
    static void TT_CAT(t_1d, SFX)(TT_CAT(tq, SFX) *q, size_t begin, size_t end, size_t s, int cubic) {       \
        size_t n = (end - begin) / s + 1;                                                                    \
        if (n <= 1) return;                                                                                  \
        FT *x = q->d;                                                                                        \
        if (!cubic || n < 5) {                                                                               \
            for (size_t i = 1; i + 1 < n; i += 2) { FT *d = x + begin + i * s; TT_CAT(tq_point, SFX)(q, d, TT_CAT(t_lin, SFX)(d[-(long)s], d[s])); } \
            if (n % 2 == 0) {                                                                                \
                FT *d = x + begin + (n - 1) * s;                                                             \
                if (n < 4) TT_CAT(tq_point, SFX)(q, d, d[-(long)s]);                                         \
                else TT_CAT(tq_point, SFX)(q, d, TT_CAT(t_lin1, SFX)(d[-3 * (long)s], d[-(long)s]));        \
            }                                                                                                \
        } else {                                                                                             \
            size_t i;                                                                                        \
            FT *d;                                                                                           \
            for (i = 3; i + 3 < n; i += 2) {                                                                 \
                d = x + begin + i * s;                                                                       \
                TT_CAT(tq_point, SFX)(q, d, TT_CAT(t_cub, SFX)(d[-3 * (long)s], d[-(long)s], d[s], d[3 * s])); \
            }                                                                                                \
            d = x + begin + s;                                                                               \
            TT_CAT(tq_point, SFX)(q, d, TT_CAT(t_q1, SFX)(d[-(long)s], d[s], d[3 * s]));                    \
            d = x + begin + i * s;                                                                           \
            TT_CAT(tq_point, SFX)(q, d, TT_CAT(t_q2, SFX)(d[-3 * (long)s], d[-(long)s], d[s]));             \
            if (n % 2 == 0) {                                                                                \
                d = x + begin + (n - 1) * s;                                                                 \
                TT_CAT(tq_point, SFX)(q, d, TT_CAT(t_q3, SFX)(d[-5 * (long)s], d[-3 * (long)s], d[-(long)s])); \
            }                                                                                                \
        }                                                                                                    \
    }                                                                                                        \
    /* foreach over a box (BlockwiseIterator.hpp:285-324) with the point rule of                             \
     * interpolation_1d_fastest_dim_first (IDH:460-526). */                                                  \
    static void TT_CAT(t_box, SFX)(TT_CAT(tq, SFX) *q, const size_t *b, const size_t *e, const size_t *st,   \
                                   const size_t *off, size_t s, size_t n, int rule, size_t bnd) {            \
        FT *x = q->d;                                                                                        \
        for (size_t i = b[0]; i < e[0]; i += st[0])                                                          \
            for (size_t j = b[1]; j < e[1]; j += st[1])                                                      \
                for (size_t k = b[2]; k < e[2]; k += st[2]) {                                                \
                    FT *d = x + i * off[0] + j * off[1] + k * off[2];                                        \
                    long S = (long)s;                                                                        \
                    FT p;                                                                                    \
                    switch (rule) {                                                                          \
                        case 0: p = TT_CAT(t_lin, SFX)(d[-S], d[S]); break;                                  \
                        case 1: p = (n < 3) ? d[-S] : TT_CAT(t_lin1, SFX)(d[-2 * S], d[-S]); break;          \
                        case 2: p = TT_CAT(t_cub, SFX)(d[-3 * S], d[-S], d[S], d[3 * S]); break;             \
                        default:                                                                             \
                            if (bnd >= 3) {                                                                  \
                                if (bnd + 3 < n) p = TT_CAT(t_cub, SFX)(d[-3 * S], d[-S], d[S], d[3 * S]);   \
                                else if (bnd + 1 < n) p = TT_CAT(t_q2, SFX)(d[-3 * S], d[-S], d[S]);         \
                                else p = TT_CAT(t_lin1, SFX)(d[-3 * S], d[-S]);                              \
                            } else {                                                                         \
                                if (bnd + 3 < n) p = TT_CAT(t_q1, SFX)(d[-S], d[S], d[3 * S]);               \
                                else if (bnd + 1 < n) p = TT_CAT(t_lin, SFX)(d[-S], d[S]);                   \
                                else p = d[-S];                                                              \
                            }                                                                                \
                    }                                                                                        \
                    TT_CAT(tq_point, SFX)(q, d, p);                                                          \
                }                                                                                            \
    }                                                                                                        \
    static void TT_CAT(t_fdf, SFX)(TT_CAT(tq, SFX) *q, const size_t *D, int dir, size_t *strides, size_t st, int cubic) { \
        size_t n = (D[dir] - 1) / st + 1;                                                                    \
        if (n <= 1) return;                                                                                  \
        size_t off[3] = {D[1] * D[2], D[2], 1}, b[3] = {0, 0, 0}, e[3] = {D[0], D[1], D[2]};                \
        size_t s = st * off[dir];                                                                            \
        off[dir] = s;                                                                                        \
        strides[dir] = 2;                                                                                    \
        if (!cubic) {                                                                                        \
            b[dir] = 1; e[dir] = n - 1;                                                                      \
            TT_CAT(t_box, SFX)(q, b, e, strides, off, s, n, 0, 0);                                           \
            if (n % 2 == 0) { b[dir] = n - 1; e[dir] = n; TT_CAT(t_box, SFX)(q, b, e, strides, off, s, n, 1, 0); } \
        } else {                                                                                             \
            b[dir] = 3; e[dir] = (n >= 3) ? n - 3 : 0;                                                       \
            TT_CAT(t_box, SFX)(q, b, e, strides, off, s, n, 2, 0);                                           \
            size_t bl[4]; int nb = 0;                                                                        \
            bl[nb++] = 1;                                                                                    \
            if (n % 2 == 1 && n > 3) bl[nb++] = n - 2;                                                       \
            if (n % 2 == 0 && n > 4) bl[nb++] = n - 3;                                                       \
            if (n % 2 == 0 && n > 2) bl[nb++] = n - 1;                                                       \
            for (int t = 0; t < nb; t++) { b[dir] = bl[t]; e[dir] = bl[t] + 1; TT_CAT(t_box, SFX)(q, b, e, strides, off, s, n, 3, bl[t]); } \
        }                                                                                                    \
    }                                                                                                        \
    /* InterpolationDecomposition<TUNING::ENABLED>::compress on one sampled block */                          \
    static void TT_CAT(t_compress, SFX)(TT_CAT(tq, SFX) *q, const size_t *D, int N, int cubic, int dir, size_t anchor_in, \
                                        double alpha, double beta, double eb) {                               \
        size_t anchor = anchor_in; int L = -1, ua = 0;                                                       \
        for (int i = 0; i < N; i++) { double lg = ceil(log2((double)D[i])); if (L < lg) L = (int)lg; if (D[i] > anchor) ua = 1; } \
        if (!ua) anchor = 0;                                                                                 \
        if (anchor > 0) { int m = (int)log2((double)anchor) + 1; if (m <= L) L = m; }                        \
        size_t Dp[3] = {1, 1, 1};                                                                            \
        for (int i = 0; i < N; i++) Dp[3 - N + i] = D[i];                                                   \
        q->eb = eb; q->ebx2 = 2 * eb; q->recip = 0.5 / eb;                                                   \
        if (anchor == 0) {                                                                                   \
            TT_CAT(tq_point, SFX)(q, q->d, (FT)0);                                                           \
        } else {                                                                                             \
            for (size_t a = 0; a < Dp[0]; a += (N > 2 ? anchor : 1))                                        \
                for (size_t bb = 0; bb < Dp[1]; bb += (N > 1 ? anchor : 1))                                 \
                    for (size_t c = 0; c < Dp[2]; c += anchor)                                               \
                        q->un[q->nun++] = q->d[a * Dp[1] * Dp[2] + bb * Dp[2] + c];                          \
            L--;                                                                                             \
        }                                                                                                    \
        for (int level = L; level > 0 && level <= L; level--) {                                              \
            double cur = eb;                                                                                 \
            if (alpha < 0) cur = (level >= 3) ? eb * 0.5 : eb;                                               \
            else if (alpha >= 1) { double r = pow(alpha, level - 1); if (r > beta) r = beta; cur = eb / r; } \
            q->eb = cur; q->ebx2 = 2 * cur; q->recip = 1.0 / q->ebx2;                                        \
            size_t st = (size_t)1 << (level - 1);                                                            \
            if (N == 1) {                                                                                    \
                TT_CAT(t_1d, SFX)(q, 0, D[0] - 1, st, cubic);                                                \
            } else if (N == 2) {                                                                             \
                int d0 = dir == 0 ? 0 : 1, d1 = 1 - d0;                                                      \
                size_t off[2] = {D[1], 1};                                                                   \
                for (size_t j = 0; j <= D[d1] - 1; j += 2 * st) {                                            \
                    size_t bo = j * off[d1];                                                                 \
                    TT_CAT(t_1d, SFX)(q, bo, bo + (D[d0] - 1) * off[d0], st * off[d0], cubic);               \
                }                                                                                            \
                for (size_t i = 0; i <= D[d0] - 1; i += st) {                                                \
                    size_t bo = i * off[d0];                                                                 \
                    TT_CAT(t_1d, SFX)(q, bo, bo + (D[d1] - 1) * off[d1], st * off[d1], cubic);               \
                }                                                                                            \
            } else {                                                                                         \
                int dims[3];                                                                                 \
                if (dir == 0) { dims[0] = 0; dims[1] = 1; dims[2] = 2; } else { dims[0] = 2; dims[1] = 1; dims[2] = 0; } \
                size_t strides[3];                                                                           \
                strides[dims[0]] = 1;                                                                        \
                for (int i = 1; i < 3; i++) strides[dims[i]] = 2 * st;                                       \
                TT_CAT(t_fdf, SFX)(q, D, dims[0], strides, st, cubic);                                       \
                for (int i = 1; i < 3; i++) {                                                                \
                    strides[dims[i - 1]] = st;                                                               \
                    TT_CAT(t_fdf, SFX)(q, D, dims[i], strides, st, cubic);                                   \
                }                                                                                            \
            }                                                                                                \
        }                                                                                                    \
    }

DEF_TUNE(float, f32, fmaf)
DEF_TUNE(double, f64, fma)

static size_t tn_lossless_size(const uint8_t *src, size_t n) {
    if (getenv("FALCONS_NOZSTD")) return n + 8;
    const fo_zstd_api *Z = fo_zstd();
    size_t cap = Z->compressBound(n);
    uint8_t *dst = (uint8_t *)malloc(cap ? cap : 1);
    size_t r = Z->compress(dst, cap, src, n, 3);
    free(dst);
    return r + 8;
}

/* interp_compress_test(_batch) (FalconsAlgoInterp.hpp:100-226): the ratio of one candidate */
static double tn_ratio(int dtype, void **blocks, size_t nb, size_t side, int N, int cubic, int dir, size_t anchor,
                       double alpha, double beta, double eb) {
    size_t D[3] = {side, side, side}, per = 1;
    for (int i = 0; i < N; i++) per *= side;
    size_t es = dtype ? 8 : 4;
    int16_t **codes = (int16_t **)malloc(nb * sizeof(void *));
    void **uns = (void **)malloc(nb * sizeof(void *));
    size_t *nq = (size_t *)malloc(nb * sizeof(size_t)), *nun = (size_t *)malloc(nb * sizeof(size_t));
    size_t total = 0;
    for (size_t b = 0; b < nb; b++) {
        void *work = malloc(per * es);
        memcpy(work, blocks[b], per * es);
        codes[b] = (int16_t *)malloc((per + 16) * 2);
        uns[b] = malloc((per + 16) * es);
        if (dtype == 0) {
            tq_f32 q = {(float *)work, codes[b], 0, (float *)uns[b], 0, 0, 0, 0};
            t_compress_f32(&q, D, N, cubic, dir, anchor, alpha, beta, eb);
            nq[b] = q.nq; nun[b] = q.nun;
        } else {
            tq_f64 q = {(double *)work, codes[b], 0, (double *)uns[b], 0, 0, 0, 0};
            t_compress_f64(&q, D, N, cubic, dir, anchor, alpha, beta, eb);
            nq[b] = q.nq; nun[b] = q.nun;
        }
        total += nq[b] * 2 + nun[b] * es;
        free(work);
    }
    uint8_t *buf = (uint8_t *)malloc(total ? total : 1), *p = buf;
    for (size_t b = 0; b < nb; b++) { memcpy(p, uns[b], nun[b] * es); p += nun[b] * es; }
    for (size_t b = 0; b < nb; b++) { memcpy(p, codes[b], nq[b] * 2); p += nq[b] * 2; }
    size_t cs = tn_lossless_size(buf, total);
    for (size_t b = 0; b < nb; b++) { free(codes[b]); free(uns[b]); }
    free(codes); free(uns); free(nq); free(nun); free(buf);
    return (double)(per * nb * es) * 1.0 / (double)cs;
}

/* ratios of the last search, in evaluation order (for tests) */
static double g_tune_ratio[16];
static int g_tune_n;

static double tn_rec(double r) {
    if (g_tune_n < 16) g_tune_ratio[g_tune_n++] = r;
    return r;
}

int fo_tune_report(double *out, int cap) {
    for (int i = 0; i < g_tune_n && i < cap; i++) out[i] = g_tune_ratio[i];
    return g_tune_n;
}

#define VAL(dtype, data, idx) ((dtype) ? ((const double *)(data))[idx] : (double)((const float *)(data))[idx])

/* Sample block side, profiling_block (Sample.hpp:5-159) and sampleBlocks (Sample.hpp:233-344) as
 * Falcons_compress_Interp_lorenzo calls them (FalconsAlgoInterp.hpp:286-324). Fills *sbs_out,
 * *per_out, *prof_out, the profiled starts (*starts_out, *nstart_out, three slots per start, the
 * field's own axis order first) and the selected block starts (*sel_out); returns the number of
 * selected blocks, or (size_t)-1 when the field is not tuned. The caller frees both arrays. */
static size_t tn_sample(const fo_config *c, int dtype, const void *data, size_t *sbs_out, size_t *per_out,
                        int *prof_out, size_t (**starts_out)[3], size_t *nstart_out, size_t (**sel_out)[3]) {
    int N = c->N;
    double rate = 0.005;
    static const size_t sbs0[4] = {4096, 128, 32, 16};
    size_t sbs = sbs0[N - 1];
    size_t shortest = c->dims[0];
    for (int i = 0; i < N; i++) if (c->dims[i] < shortest) shortest = c->dims[i];
    while (sbs >= shortest) sbs /= 2;
    while (sbs >= 16 && (pow((double)(sbs + 1), N) / (double)c->num) > 1.5 * rate) sbs /= 2;
    if (sbs < 8) sbs = 8;
    int to_tune = pow((double)(sbs + 1), N) <= 0.05 * (double)c->num;
    for (int i = 0; i < N; i++) if (c->dims[i] < sbs) to_tune = 0;
    if (!to_tune) return (size_t)-1;
    size_t per = (size_t)pow((double)(sbs + 1), N);
    size_t es = dtype ? 8 : 4;
    size_t D[3] = {1, 1, 1};
    for (int i = 0; i < N; i++) D[3 - N + i] = c->dims[i];
    /* profiling_block (Sample.hpp:5-159): blocks whose strided samples vary by more than eb */
    size_t stride = sbs / 4;
    if (stride == 0) stride = sbs;
    size_t nstart = 0, cap = 1024;
    size_t (*starts)[3] = malloc(cap * sizeof *starts);
    {
        size_t n_i = (D[3 - N] - sbs + sbs - 1) / sbs;
        size_t lim1 = (N >= 2) ? D[3 - N + 1] - sbs : 1, lim2 = (N >= 3) ? D[2] - sbs : 1;
        size_t st1 = (N >= 2) ? sbs : 1, st2 = (N >= 3) ? sbs : 1;
        for (size_t vi = 0; vi < n_i; vi++) {
            size_t i = vi * sbs;
            for (size_t j = 0; j < lim1; j += st1)
                for (size_t k = 0; k < lim2; k += st2) {
                    size_t a[3] = {i, j, k};
                    size_t s0 = 0;
                    if (N == 3) s0 = a[0] * D[1] * D[2] + a[1] * D[2] + a[2];
                    else if (N == 2) s0 = a[0] * D[2] + a[1];
                    else s0 = a[0];
                    double mn = VAL(dtype, data, s0), mx = mn;
                    size_t r1 = N >= 2 ? sbs : 0, r2 = N >= 3 ? sbs : 0;
                    for (size_t ii = 0; ii <= sbs; ii += stride)
                        for (size_t jj = 0; jj <= r1; jj += stride)
                            for (size_t kk = 0; kk <= r2; kk += stride) {
                                size_t idx;
                                if (N == 3) idx = s0 + ii * D[1] * D[2] + jj * D[2] + kk;
                                else if (N == 2) idx = s0 + ii * D[2] + jj;
                                else idx = s0 + ii;
                                double v = VAL(dtype, data, idx);
                                if (v < mn) mn = v;
                                else if (v > mx) mx = v;
                            }
                    /* comparisons above happen in the element type in the reference; for float the
                     * promotion to double preserves order and equality */
                    double range = dtype ? (mx - mn) : (double)((float)mx - (float)mn);
                    if (range > c->abs_eb) {
                        if (nstart == cap) { cap *= 2; starts = realloc(starts, cap * sizeof *starts); }
                        starts[nstart][0] = i; starts[nstart][1] = j; starts[nstart][2] = k;
                        nstart++;
                    }
                }
        }
    }
    int profiling = (double)(nstart * per) >= 0.5 * rate * (double)c->num;
    /* sampleBlocks (Sample.hpp:233-344) */
    size_t totalblock = 1;
    for (int i = 0; i < N; i++) totalblock *= (size_t)(int)((c->dims[i] - 1) / sbs);
    size_t nsel = 0, selcap = 1024;
    size_t (*sel)[3] = malloc(selcap * sizeof *sel);
    if (profiling) {
        size_t sstride = (size_t)((double)nstart / ((double)totalblock * rate));
        if (sstride == 0) sstride = 1;
        size_t n_sel = (nstart + sstride - 1) / sstride;
        if (n_sel > selcap) { selcap = n_sel; sel = realloc(sel, selcap * sizeof *sel); }
        for (size_t si = 0; si < n_sel; si++) { memcpy(sel[si], starts[si * sstride], sizeof sel[0]); }
        nsel = n_sel;
    } else {
        size_t sstride = (size_t)(1.0 / rate);
        if (sstride == 0) sstride = 1;
        size_t idx = 0;
        size_t l0 = D[3 - N] - sbs, l1 = (N >= 2) ? D[3 - N + 1] - sbs : 1, l2 = (N >= 3) ? D[2] - sbs : 1;
        for (size_t x = 0; x < l0; x += sbs)
            for (size_t y = 0; y < l1; y += (N >= 2 ? sbs : 1))
                for (size_t z = 0; z < l2; z += (N >= 3 ? sbs : 1)) {
                    if (idx % sstride == 0) {
                        if (nsel == selcap) { selcap *= 2; sel = realloc(sel, selcap * sizeof *sel); }
                        sel[nsel][0] = x; sel[nsel][1] = y; sel[nsel][2] = z;
                        nsel++;
                    }
                    idx++;
                }
    }
    *sbs_out = sbs;
    *per_out = per;
    *prof_out = profiling;
    *starts_out = starts;
    *nstart_out = nstart;
    *sel_out = sel;
    (void)es;
    return nsel;
}

long fo_tune_sample(const fo_config *c, int dtype, const void *data, size_t *sbs, int *profiling, size_t *starts,
                    size_t starts_cap, size_t *nstarts, size_t *sel, size_t sel_cap) {
    size_t per = 0, nst = 0;
    size_t(*st)[3] = NULL, (*sl)[3] = NULL;
    size_t n = tn_sample(c, dtype, data, sbs, &per, profiling, &st, &nst, &sl);
    if (n == (size_t)-1) return -1;
    *nstarts = nst;
    for (size_t i = 0; i < nst && i < starts_cap; i++) memcpy(starts + 3 * i, st[i], 3 * sizeof(size_t));
    for (size_t i = 0; i < n && i < sel_cap; i++) memcpy(sel + 3 * i, sl[i], 3 * sizeof(size_t));
    free(st);
    free(sl);
    return (long)n;
}

int fo_tune_interp(fo_config *c, int dtype, const void *data, int omp_threads) {
    int N = c->N;
    g_tune_n = 0;
    if (c->interp_anchor_stride < 0) {
        static const int def[4] = {4096, 128, 32, 16};
        c->interp_anchor_stride = def[N - 1];
    }
    size_t sbs, per, nstart;
    int profiling;
    size_t(*starts)[3], (*sel)[3];
    size_t nsel = tn_sample(c, dtype, data, &sbs, &per, &profiling, &starts, &nstart, &sel);
    if (nsel == (size_t)-1) return 0;
    free(starts);
    size_t es = dtype ? 8 : 4;
    size_t D[3] = {1, 1, 1};
    for (int i = 0; i < N; i++) D[3 - N + i] = c->dims[i];
    if (nsel * per == 0 || (double)(nsel * per) >= (double)c->num * 0.2) { free(sel); return 0; }
    void **blocks = (void **)malloc(nsel * sizeof(void *));
    size_t side = sbs + 1;
    for (size_t b = 0; b < nsel; b++) {
        blocks[b] = malloc(per * es);
        size_t o = 0;
        for (size_t i = 0; i < side; i++)
            for (size_t j = 0; j < (N >= 2 ? side : 1); j++)
                for (size_t k = 0; k < (N >= 3 ? side : 1); k++) {
                    size_t idx;
                    if (N == 3) idx = (i + sel[b][0]) * D[1] * D[2] + (j + sel[b][1]) * D[2] + k + sel[b][2];
                    else if (N == 2) idx = (i + sel[b][0]) * D[2] + (j + sel[b][1]);
                    else idx = i + sel[b][0];
                    memcpy((uint8_t *)blocks[b] + o * es, (const uint8_t *)data + idx * es, es);
                    o++;
                }
    }
    free(sel);
    double best = 0;
    size_t anchor = (size_t)c->interp_anchor_stride;
    int reverse = (N == 3) ? 5 : (N == 2 ? 1 : 0);
    c->interp_direction = 0;
    if (N <= 2) { c->interp_alpha = 1.0; c->interp_beta = 1.0; }
    else { c->interp_alpha = 1.25; c->interp_beta = 2.0; }
    if (N >= 3) {
        for (int i = 0; i < 4; i++) {
            int cub = (i / 2 == 0) ? 0 : 1, dir = (i % 2 == 0) ? 0 : reverse;
            double r = tn_rec(tn_ratio(dtype, blocks, nsel, side, N, cub, dir, anchor, c->interp_alpha, c->interp_beta, c->abs_eb));
            if (r > best) { best = r; c->interp_algo = cub; c->interp_direction = dir; }
        }
    } else if (omp_threads >= 6) {
        static const double al[2] = {1.0, 1.25}, be[2] = {1.0, 2.0};
        for (int i = 0; i < 8; i++) {
            int ad = i % 4, ab = i / 4;
            int cub = (ad / 2 == 0) ? 0 : 1, dir = (ad % 2 == 0) ? 0 : reverse;
            double r = tn_rec(tn_ratio(dtype, blocks, nsel, side, N, cub, dir, anchor, al[ab], be[ab], c->abs_eb));
            if (r > best) { best = r; c->interp_algo = cub; c->interp_direction = dir; c->interp_alpha = al[ab]; c->interp_beta = be[ab]; }
        }
    } else {
        for (int i = 0; i < 4; i++) {
            int cub = (i / 2 == 0) ? 0 : 1, dir = (i % 2 == 0) ? 0 : reverse;
            double r = tn_rec(tn_ratio(dtype, blocks, nsel, side, N, cub, dir, anchor, c->interp_alpha, c->interp_beta, c->abs_eb));
            if (r > best) { best = r; c->interp_algo = cub; c->interp_direction = dir; }
        }
        static const double al[3] = {1.0, 1.25, 1.5}, be[3] = {1.0, 2.0, 2.5};
        double abr[3];
        for (int i = 1; i < 3; i++)
            abr[i] = tn_rec(tn_ratio(dtype, blocks, nsel, side, N, c->interp_algo, c->interp_direction, anchor, al[i], be[i], c->abs_eb));
        for (int i = 1; i < 3; i++)
            if (abr[i] > best * 1.02) { best = abr[i]; c->interp_alpha = al[i]; c->interp_beta = be[i]; }
    }
    for (size_t b = 0; b < nsel; b++) free(blocks[b]);
    free(blocks);
    /* 1-D data below ratio 50 goes on to the Lorenzo test (FalconsAlgoInterp.hpp:479-497),
     * which belongs to a different algorithm */
    if (N == 1 && best < 50) return 2;
    return 1;
}
