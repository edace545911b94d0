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

static size_t tn_lossless_size(const uint8_t *src, size_t n) {
    if (getenv("FALCONS_NOZSTD")) return n + 8;
    const fo_zstd_api *Z = fo_zstd();
    size_t cap = Z->compressBound(n);
    uint8_t *dst = (uint8_t *)malloc(cap ? cap : 1);
    size_t r = Z->compress(dst, cap, src, n, 3);
    free(dst);
    return r + 8;
}

/* interp_compress_test(_batch) (FalconsAlgoInterp.hpp:100-226): the ratio of one candidate. Each
 * block goes through InterpolationDecomposition<TUNING::ENABLED>::compress, whose traversal is the
 * generic one (oracle_generic.c); the buffer is [all unpredictables | all codes]. */
static double tn_ratio(int dtype, void **blocks, size_t nb, size_t side, int N, int cubic, int dir, size_t anchor,
                       double alpha, double beta, double eb) {
    size_t D[4] = {side, side, side, side}, per = 1;
    for (int i = 0; i < N; i++) per *= side;
    size_t es = fo_elem_size(dtype);
    int16_t **codes = (int16_t **)malloc(nb * sizeof(void *));
    void **uns = (void **)malloc(nb * sizeof(void *));
    size_t *nq = (size_t *)malloc(nb * sizeof(size_t)), *nun = (size_t *)malloc(nb * sizeof(size_t));
    size_t total = 0;
    for (size_t b = 0; b < nb; b++) {
        void *work = malloc(per * es);
        memcpy(work, blocks[b], per * es);
        codes[b] = (int16_t *)malloc((per + 16) * 2);
        uns[b] = malloc((per + 16) * es);
        size_t anchor_used;
        fo_gen_compress(dtype, N, D, cubic, dir, (long)anchor, alpha, beta, eb, work, codes[b], &nq[b], uns[b],
                        per + 16, &nun[b], &anchor_used);
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

/* profiling_block's test of one candidate block (Sample.hpp:5-159): min and max over the points every
 * `stride` along each axis, compared in the element type, the range against eb. */
#define TN_PROF(T)                                                                                   \
    {                                                                                                \
        const T *d = (const T *)data;                                                                \
        T mn = d[s0], mx = mn;                                                                       \
        for (size_t i0 = 0; i0 <= lim[0]; i0 += stride)                                              \
            for (size_t i1 = 0; i1 <= lim[1]; i1 += stride)                                          \
                for (size_t i2 = 0; i2 <= lim[2]; i2 += stride)                                      \
                    for (size_t i3 = 0; i3 <= lim[3]; i3 += stride) {                                \
                        T v = d[s0 + i0 * off[0] + i1 * off[1] + i2 * off[2] + i3 * off[3]];         \
                        if (v < mn) mn = v;                                                          \
                        else if (v > mx) mx = v;                                                     \
                    }                                                                                \
        return (double)(mx - mn) > eb;                                                               \
    }
static int tn_profile(int dtype, const void *data, size_t s0, const size_t *off, const size_t *lim, size_t stride,
                      double eb) {
    switch (dtype) {
        case 0: TN_PROF(float)
        case 1: TN_PROF(double)
        case 2: TN_PROF(uint8_t)
        case 3: TN_PROF(int8_t)
        case 4: TN_PROF(uint16_t)
        case 5: TN_PROF(int16_t)
        case 6: TN_PROF(uint32_t)
        case 7: TN_PROF(int32_t)
        case 8: TN_PROF(uint64_t)
        default: TN_PROF(int64_t)
    }
}
#undef TN_PROF

/* Sample block side, profiling_block (Sample.hpp:5-159) and sampleBlocks (Sample.hpp:233-344) as
 * Falcons_compress_Interp_lorenzo calls them (FalconsAlgoInterp.hpp:286-324). The field's axes sit in
 * the last N of four slots. Fills *sbs_out, *per_out, *prof_out, the profiled starts (*starts_out,
 * *nstart_out, four slots per start) and the selected block starts (*sel_out); returns the number of
 * selected blocks, or (size_t)-1 when the field is not tuned. The caller frees both arrays. */
static size_t tn_sample(const fo_config *c, int dtype, const void *data, size_t *sbs_out, size_t *per_out,
                        int *prof_out, size_t (**starts_out)[4], size_t *nstart_out, size_t (**sel_out)[4]) {
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
    size_t D[4] = {1, 1, 1, 1}, off[4], ns[4], lim[4], nstart = 1;
    for (int i = 0; i < N; i++) D[4 - N + i] = c->dims[i];
    off[3] = 1;
    for (int a = 2; a >= 0; a--) off[a] = off[a + 1] * D[a + 1];
    for (int a = 0; a < 4; a++) {
        const int real = a >= 4 - N;
        ns[a] = real ? (D[a] - sbs + sbs - 1) / sbs : 1;   /* starts 0, sbs, ... below D - sbs */
        lim[a] = real ? sbs : 0;
        nstart *= ns[a];
    }
    size_t stride = sbs / 4;
    if (stride == 0) stride = sbs;
    size_t(*starts)[4] = malloc((nstart + 1) * sizeof *starts), nst = 0;
    for (size_t t = 0; t < nstart; t++) {
        size_t v[4], r = t, s0 = 0;
        for (int a = 3; a >= 0; a--) { v[a] = (r % ns[a]) * (a >= 4 - N ? sbs : 0); r /= ns[a]; s0 += v[a] * off[a]; }
        if (tn_profile(dtype, data, s0, off, lim, stride, c->abs_eb)) { memcpy(starts[nst], v, sizeof v); nst++; }
    }
    int profiling = (double)(nst * per) >= 0.5 * rate * (double)c->num;
    size_t totalblock = 1;
    for (int i = 0; i < N; i++) totalblock *= (size_t)(int)((c->dims[i] - 1) / sbs);
    size_t(*sel)[4] = malloc((nstart + 1) * sizeof *sel), nsel = 0;
    if (profiling) {
        size_t sstride = (size_t)((double)nst / ((double)totalblock * rate));
        if (sstride == 0) sstride = 1;
        size_t n_sel = (nst + sstride - 1) / sstride;
        for (size_t si = 0; si < n_sel; si++) memcpy(sel[nsel++], starts[si * sstride], sizeof sel[0]);
    } else {
        size_t sstride = (size_t)(1.0 / rate);
        if (sstride == 0) sstride = 1;
        for (size_t t = 0; t < nstart; t++) {
            if (t % sstride) continue;
            size_t r = t;
            for (int a = 3; a >= 0; a--) { sel[nsel][a] = (r % ns[a]) * (a >= 4 - N ? sbs : 0); r /= ns[a]; }
            nsel++;
        }
    }
    *sbs_out = sbs;
    *per_out = per;
    *prof_out = profiling;
    *starts_out = starts;
    *nstart_out = nst;
    *sel_out = sel;
    return nsel;
}

long fo_tune_sample(const fo_config *c, int dtype, const void *data, size_t *sbs, int *profiling, size_t *starts,
                    size_t starts_cap, size_t *nstarts, size_t *sel, size_t sel_cap) {
    size_t per = 0, nst = 0;
    size_t(*st)[4] = NULL, (*sl)[4] = NULL;
    size_t n = tn_sample(c, dtype, data, sbs, &per, profiling, &st, &nst, &sl);
    if (n == (size_t)-1) return -1;
    *nstarts = nst;
    /* the field's own axes first */
    const int o = 4 - c->N;
    for (size_t i = 0; i < nst && i < starts_cap; i++)
        for (int a = 0; a < 4; a++) starts[4 * i + a] = a < c->N ? st[i][o + a] : 0;
    for (size_t i = 0; i < n && i < sel_cap; i++)
        for (int a = 0; a < 4; a++) sel[4 * i + a] = a < c->N ? sl[i][o + a] : 0;
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
    size_t(*starts)[4], (*sel)[4];
    size_t nsel = tn_sample(c, dtype, data, &sbs, &per, &profiling, &starts, &nstart, &sel);
    if (nsel == (size_t)-1) return 0;
    free(starts);
    size_t es = fo_elem_size(dtype);
    size_t D[4] = {1, 1, 1, 1};
    for (int i = 0; i < N; i++) D[4 - N + i] = c->dims[i];
    if (nsel * per == 0 || (double)(nsel * per) >= (double)c->num * 0.2) { free(sel); return 0; }
    void **blocks = (void **)malloc(nsel * sizeof(void *));
    size_t side = sbs + 1;
    for (size_t b = 0; b < nsel; b++) {
        blocks[b] = malloc(per * es);
        size_t o = 0, ext[4];
        for (int a = 0; a < 4; a++) ext[a] = a >= 4 - N ? side : 1;
        for (size_t i0 = 0; i0 < ext[0]; i0++)
            for (size_t i1 = 0; i1 < ext[1]; i1++)
                for (size_t i2 = 0; i2 < ext[2]; i2++)
                    for (size_t i3 = 0; i3 < ext[3]; i3++) {
                        size_t idx = (((i0 + sel[b][0]) * D[1] + i1 + sel[b][1]) * D[2] + i2 + sel[b][2]) * D[3] +
                                     i3 + sel[b][3];
                        memcpy((uint8_t *)blocks[b] + o * es, (const uint8_t *)data + idx * es, es);
                        o++;
                    }
    }
    free(sel);
    double best = 0;
    size_t anchor = (size_t)c->interp_anchor_stride;
    int reverse = 1; /* factorial(N) - 1 (FalconsAlgoInterp.hpp:355) */
    for (int i = 2; i <= N; i++) reverse *= i;
    reverse -= 1;
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
