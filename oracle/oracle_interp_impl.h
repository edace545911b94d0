/* Interpolation decomposition restatement, instantiated once per element type by
 * oracle_interp.c with FT = float / double and SFX = f32 / f64.
 * Follows include/Falcons/decomposition/InterpolationDecomposition.hpp (IDH) and the AVX2 section
 * of Interpolation_quantizer.inl (.inl) loop nest by loop nest, so that the code order, the
 * unpredictable order and the per-point compare precision are those of the serial AVX2 build. */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* Interpolators.hpp:48-117 (fp contraction is off for this file; fused steps are explicit fma) */
static inline FT FN(i_lin)(FT a, FT b) { return (a + b) / 2; }
static inline FT FN(i_lin1)(FT a, FT b) { return FN(fo_fma)(-(FT)0.5, a, (FT)1.5 * b); }
static inline FT FN(i_q1)(FT a, FT b, FT c) {
    FT t = FN(fo_fma)((FT)6, b, -c);
    t = FN(fo_fma)((FT)3, a, t);
    return t * (FT)0.125;
}
static inline FT FN(i_q2)(FT a, FT b, FT c) {
    FT t = FN(fo_fma)((FT)6, b, -a);
    t = FN(fo_fma)((FT)3, c, t);
    return t * (FT)0.125;
}
static inline FT FN(i_cub)(FT a, FT b, FT c, FT d) {
    FT t = (b + c) * (FT)9;
    t = t - a;
    t = t - d;
    return t * (FT)0.0625;
}

typedef struct {
    FT *data;
    int16_t *codes;
    size_t nq;
    FT *unpred;
    size_t nun, un_cap;
    int decomp;
    /* current level's quantizer constants (LinearQuantizer.hpp:39-43, IDH:2301-2311) */
    double eb, ebx2, recip;
    float ebf;
    const FT *unpred_in;
    size_t unpred_pos;
    int err;
} FN(fo_ctx);

static void FN(push_unpred)(FN(fo_ctx) *c, FT v) {
    if (c->nun < c->un_cap) c->unpred[c->nun] = v;
    c->nun++;
}

/* One target point. vclass != 0: AVX2 lane compare (f32 error against f32 bound, .inl:1200-1239);
 * vclass == 0: LinearQuantizer::quantize_and_overwrite (double compare, LinearQuantizer.hpp:49-82).
 * For double elements both classes are the same arithmetic. */
static void FN(point)(FN(fo_ctx) *c, size_t idx, FT pred, int vclass, int skip) {
    FT *d = &c->data[idx];
    if (c->decomp) {
        int q = c->codes[c->nq++];
        if (q == -32768) {
            *d = c->unpred_in[c->unpred_pos++];
        } else {
            *d = (FT)fma((double)q, c->ebx2, (double)pred);
        }
        return;
    }
    FT ori = *d;
    FT diff = ori - pred;
    double qd = nearbyint((double)diff * c->recip);
    int code;
    FT dec = 0;
    int ok = (qd > -32768.0 && qd < 32768.0);
    if (ok) {
        dec = (FT)fma(qd, c->ebx2, (double)pred);
        FT err = dec - ori;
        if (vclass && fo_g_avx2 && sizeof(FT) == 4) {
            float e = (float)err, ebf = c->ebf;
            ok = (e >= -ebf && e <= ebf);
        } else {
            ok = ((double)err >= -c->eb && (double)err <= c->eb);
        }
    }
    if (ok) {
        code = (int)qd;
        if (!skip) *d = dec;
    } else {
        code = -32768;
        FN(push_unpred)(c, ori);
    }
    c->codes[c->nq++] = (int16_t)code;
}

/* interp_linear_and_quantize_1D_line (.inl:32-180): one line along the pass axis */
static void FN(line_linear)(FN(fo_ctx) *c, size_t base, size_t n, size_t sp, int skip) {
    if (n <= 1) return;
    size_t odd_len = n / 2, even_len = n - odd_len;
    FT *x = c->data;
    size_t vec = (sizeof(FT) == 4) ? 8 : 4, i = 0;
    for (; i + vec < odd_len; i += vec) {
        FT pred[8];
        for (size_t j = 0; j < vec; j++) {
            size_t e = (i + j) * 2;
            pred[j] = (x[base + e * sp] + x[base + (e + 2) * sp]) * (FT)0.5;
        }
        for (size_t j = 0; j < vec; j++) FN(point)(c, base + (2 * (i + j) + 1) * sp, pred[j], 1, skip);
    }
    for (; i + 1 < odd_len; i++)
        FN(point)(c, base + (2 * i + 1) * sp, FN(i_lin)(x[base + 2 * i * sp], x[base + (2 * i + 2) * sp]), 0, skip);
    FT pe = (n < 3) ? x[base + 2 * (even_len - 1) * sp]
                    : FN(i_lin1)(x[base + 2 * (even_len - 2) * sp], x[base + 2 * (even_len - 1) * sp]);
    FN(point)(c, base + (2 * odd_len - 1) * sp, pe, 0, skip);
}

/* interp_cubic_and_quantize_1D_line (.inl:184-377) */
static void FN(line_cubic)(FN(fo_ctx) *c, size_t base, size_t n, size_t sp, int skip) {
    if (n <= 1) return;
    size_t odd_len = n / 2, even_len = n - odd_len;
    FT *x = c->data;
#define X(e) x[base + (size_t)(e) * sp]
    FT pf;
    if (even_len < 2) pf = X(0);
    else if (even_len < 3) pf = FN(i_lin)(X(0), X(2));
    else pf = FN(i_q1)(X(0), X(2), X(4));
    FN(point)(c, base + sp, pf, 0, skip);
    size_t vec = (sizeof(FT) == 4) ? 8 : 4, i = 0;
    for (; i + vec + 2 < even_len; i += vec) {
        FT pred[8];
        for (size_t j = 0; j < vec; j++) {
            size_t e = (i + j) * 2;
            pred[j] = ((((X(e + 2) + X(e + 4)) * (FT)9) - X(e)) - X(e + 6)) * ((FT)1 / (FT)16);
        }
        for (size_t j = 0; j < vec; j++) FN(point)(c, base + (2 * (i + j) + 3) * sp, pred[j], 1, skip);
    }
    for (; i + 3 < even_len; i++)
        FN(point)(c, base + (2 * i + 3) * sp, FN(i_cub)(X(2 * i), X(2 * i + 2), X(2 * i + 4), X(2 * i + 6)), 0, skip);
    if (odd_len > 1) {
        if (odd_len < even_len) {
            FT ep = FN(i_q2)(X(2 * (even_len - 3)), X(2 * (even_len - 2)), X(2 * (even_len - 1)));
            FN(point)(c, base + (2 * odd_len - 1) * sp, ep, 0, skip);
        } else {
            if (odd_len > 2) {
                FT ep = FN(i_q2)(X(2 * (even_len - 3)), X(2 * (even_len - 2)), X(2 * (even_len - 1)));
                FN(point)(c, base + (2 * odd_len - 3) * sp, ep, 0, skip);
            }
            FT ep = FN(i_lin1)(X(2 * (even_len - 2)), X(2 * (even_len - 1)));
            FN(point)(c, base + (2 * odd_len - 1) * sp, ep, 0, skip);
        }
    }
#undef X
}

/* A row of the row-vector kernels (.inl:787-1109): all predictions of the row are formed from the
 * predictor rows first, then the row's points are quantized in increasing fastest-dim order.
 * kind: 0 linear(a,b) 1 cubic(a,b,c,d) 2 equal(a) 3 linear1(a,b) 4 quad1(a,b,c) 5 quad2(a,b,c).
 * rows[] hold element offsets of the predictor rows; tgt is the target row offset. */
static void FN(vrow)(FN(fo_ctx) *c, int kind, const size_t *rows, size_t tgt, size_t len, size_t sk, int skip) {
    FT *x = c->data;
    for (size_t k = 0; k < len; k++) {
        size_t o = k * sk;
        FT p;
        switch (kind) {
            case 0: p = (x[rows[0] + o] + x[rows[1] + o]) * (FT)0.5; break;
            case 1: p = ((((x[rows[1] + o] + x[rows[2] + o]) * (FT)9) - x[rows[0] + o]) - x[rows[3] + o]) * ((FT)1 / (FT)16); break;
            case 2: p = x[rows[0] + o]; break;
            case 3: p = FN(fo_fma)(-(FT)0.5, x[rows[0] + o], x[rows[1] + o] * (FT)1.5); break;
            case 4: p = FN(fo_fma)(x[rows[0] + o], (FT)3, FN(fo_fma)(x[rows[1] + o], (FT)6, -x[rows[2] + o])) * (FT)0.125; break;
            default: p = FN(fo_fma)(x[rows[2] + o], (FT)3, FN(fo_fma)(x[rows[1] + o], (FT)6, -x[rows[0] + o])) * (FT)0.125; break;
        }
        FN(point)(c, tgt + o, p, 1, skip);
    }
}

/* The row passes: interpolation_1d_simd_3d_z (IDH:819-1135), _3d_y (IDH:1138-1375) and
 * _2d_y (IDH:604-768). The pass axis has n grid points spaced sp elements apart; the rows run
 * along the fastest axis (len points, sk apart); no = number of outer rows (3D: j for the z pass,
 * i for the y pass; 2D: 1), so = their element spacing. The even-n cubic boundary plane is
 * coded after all outer rows by the scalar quantizer in 3D, and by the vector l1 kernel at the end
 * of the only outer row in 2D. */
static void FN(pass_rows)(FN(fo_ctx) *c, int cubic, size_t n, size_t sp, size_t no, size_t so, size_t len,
                          size_t sk, int skip, int is2d) {
    if (n <= 1) return;
    size_t r[4];
    if (!cubic) {
        size_t lls = skip ? 3 * sp : 2 * sp;
        for (size_t o = 0; o < no; o++) {
            size_t ob = o * so;
            for (size_t g = 1; g + 1 < n; g += 2) {
                r[0] = ob + (g - 1) * sp; r[1] = ob + (g + 1) * sp;
                FN(vrow)(c, 0, r, ob + g * sp, len, sk, skip);
            }
            if (n % 2 == 0) {
                size_t t = ob + (n - 1) * sp;
                if (n < 3) {
                    r[0] = t - sp;
                    FN(vrow)(c, 2, r, t, len, sk, skip);
                } else {
                    r[0] = t - lls; r[1] = t - sp;
                    FN(vrow)(c, 3, r, t, len, sk, skip);
                }
            }
        }
        return;
    }
    for (size_t o = 0; o < no; o++) {
        size_t ob = o * so;
        size_t t1 = ob + sp;
        if (n >= 5) { r[0] = ob; r[1] = ob + 2 * sp; r[2] = ob + 4 * sp; FN(vrow)(c, 4, r, t1, len, sk, skip); }
        else if (n >= 3) { r[0] = ob; r[1] = ob + 2 * sp; FN(vrow)(c, 0, r, t1, len, sk, skip); }
        else { r[0] = ob; FN(vrow)(c, 2, r, t1, len, sk, skip); }
        for (size_t g = 3; g + 3 < n; g += 2) {
            r[0] = ob + (g - 3) * sp; r[1] = ob + (g - 1) * sp; r[2] = ob + (g + 1) * sp; r[3] = ob + (g + 3) * sp;
            FN(vrow)(c, 1, r, ob + g * sp, len, sk, skip);
        }
        if (n % 2 == 1) {
            if (n > 3) {
                r[0] = ob + (n - 5) * sp; r[1] = ob + (n - 3) * sp; r[2] = ob + (n - 1) * sp;
                FN(vrow)(c, 5, r, ob + (n - 2) * sp, len, sk, skip);
            }
        } else {
            if (n > 4) {
                r[0] = ob + (n - 6) * sp; r[1] = ob + (n - 4) * sp; r[2] = ob + (n - 2) * sp;
                FN(vrow)(c, 5, r, ob + (n - 3) * sp, len, sk, skip);
            }
            if (is2d && n > 2) {
                r[0] = ob + (n - 4) * sp; r[1] = ob + (n - 2) * sp;
                FN(vrow)(c, 3, r, ob + (n - 1) * sp, len, sk, skip);
            }
        }
    }
    if (!is2d && n % 2 == 0 && n > 2) {
        /* foreach over the boundary plane (IDH:1095-1132 / 1335-1372): boundary = n-1 >= 3 and
         * boundary + 1 == n, so the formula is interp_linear1(d - 3s, d - s), scalar quantizer */
        FT *x = c->data;
        for (size_t o = 0; o < no; o++)
            for (size_t k = 0; k < len; k++) {
                size_t t = o * so + (n - 1) * sp + k * sk;
                FN(point)(c, t, FN(i_lin1)(x[t - 3 * sp], x[t - sp]), 0, skip);
            }
    }
}

/* interpolation_1d_simd_3d_x / _2d_x / interpolation_1d_simd: lines over two outer loops */
static void FN(pass_lines)(FN(fo_ctx) *c, int cubic, size_t n, size_t sp, size_t n1, size_t s1, size_t n2,
                           size_t s2, int skip) {
    if (n <= 1) return;
    for (size_t a = 0; a < n1; a++)
        for (size_t b = 0; b < n2; b++) {
            size_t base = a * s1 + b * s2;
            if (cubic) FN(line_cubic)(c, base, n, sp, skip);
            else FN(line_linear)(c, base, n, sp, skip);
        }
}

/* InterpolationDecomposition::interpolation() dispatch for one level (IDH:2294-2510) */
static int FN(level)(FN(fo_ctx) *c, const size_t *D, int N, int cubic, int dir, size_t st, int last) {
    if (N == 1) {
        FN(pass_lines)(c, cubic, (D[0] - 1) / st + 1, st, 1, 0, 1, 0, last);
        return 0;
    }
    if (N == 2) {
        size_t D0 = D[0], D1 = D[1];
        if (dir == 0) { /* 2d_y then 2d_x */
            FN(pass_rows)(c, cubic, (D0 - 1) / st + 1, st * D1, 1, 0, (D1 - 1) / (2 * st) + 1, 2 * st, 0, 1);
            FN(pass_lines)(c, cubic, (D1 - 1) / st + 1, st, (D0 - 1) / st + 1, st * D1, 1, 0, last);
        } else {        /* 2d_x then 2d_y */
            FN(pass_lines)(c, cubic, (D1 - 1) / st + 1, st, (D0 - 1) / (2 * st) + 1, 2 * st * D1, 1, 0, 0);
            FN(pass_rows)(c, cubic, (D0 - 1) / st + 1, st * D1, 1, 0, (D1 - 1) / st + 1, st, last, 1);
        }
        return 0;
    }
    size_t Dz = D[0], Dy = D[1], Dx = D[2], Sz = Dy * Dx, Sy = Dx;
    if (dir == 0) { /* z, y, x */
        FN(pass_rows)(c, cubic, (Dz - 1) / st + 1, st * Sz, (Dy - 1) / (2 * st) + 1, 2 * st * Sy,
                      (Dx - 1) / (2 * st) + 1, 2 * st, 0, 0);
        FN(pass_rows)(c, cubic, (Dy - 1) / st + 1, st * Sy, (Dz - 1) / st + 1, st * Sz, (Dx - 1) / (2 * st) + 1,
                      2 * st, 0, 0);
        FN(pass_lines)(c, cubic, (Dx - 1) / st + 1, st, (Dz - 1) / st + 1, st * Sz, (Dy - 1) / st + 1, st * Sy, last);
    } else if (dir == 5) { /* x, y, z */
        FN(pass_lines)(c, cubic, (Dx - 1) / st + 1, st, (Dz - 1) / (2 * st) + 1, 2 * st * Sz, (Dy - 1) / (2 * st) + 1,
                       2 * st * Sy, 0);
        FN(pass_rows)(c, cubic, (Dy - 1) / st + 1, st * Sy, (Dz - 1) / (2 * st) + 1, 2 * st * Sz, (Dx - 1) / st + 1,
                      st, 0, 0);
        FN(pass_rows)(c, cubic, (Dz - 1) / st + 1, st * Sz, (Dy - 1) / st + 1, st * Sy, (Dx - 1) / st + 1, st, last, 0);
    } else {
        return -1;
    }
    return 0;
}

/* InterpolationDecomposition::init (IDH:286-339): number of levels and whether anchors are used */
static void FN(plan)(const size_t *D, int N, size_t *anchor, int *levels) {
    int L = -1, use_anchor = 0;
    for (int i = 0; i < N; i++) {
        double lg = ceil(log2((double)D[i]));
        if (L < lg) L = (int)lg;
        if (D[i] > *anchor) use_anchor = 1;
    }
    if (!use_anchor) *anchor = 0;
    if (*anchor > 0) {
        int m = (int)log2((double)*anchor) + 1;
        if (m <= L) L = m;
    }
    *levels = L;
}

/* compress (IDH:131-239) / decompress (IDH:48-128) */
static int FN(run)(FN(fo_ctx) *c, const size_t *D, int N, int cubic, int dir, size_t anchor_in, double alpha,
                   double beta, double eb, size_t *anchor_out) {
    size_t anchor = anchor_in;
    int L;
    FN(plan)(D, N, &anchor, &L);
    *anchor_out = anchor;
    size_t off[3] = {1, 1, 1};
    if (N == 3) { off[0] = D[1] * D[2]; off[1] = D[2]; off[2] = 1; }
    else if (N == 2) { off[0] = D[1]; off[1] = 1; }
    if (anchor == 0) {
        /* first point against prediction 0 with the constructor's reciprocal 0.5/eb */
        c->eb = eb; c->ebx2 = 2 * eb; c->recip = 0.5 / eb; c->ebf = (float)eb;
        FN(point)(c, 0, (FT)0, 0, 0);
    } else {
        /* anchor grid, row-major, kept losslessly (IDH:341-359) */
        size_t cnt[3] = {1, 1, 1};
        for (int i = 0; i < N; i++) cnt[i] = (D[i] - 1) / anchor + 1;
        for (size_t a = 0; a < cnt[0]; a++)
            for (size_t b = 0; b < (N > 1 ? cnt[1] : 1); b++)
                for (size_t d = 0; d < (N > 2 ? cnt[2] : 1); d++) {
                    size_t idx = a * anchor * off[0] + (N > 1 ? b * anchor * off[1] : 0) + (N > 2 ? d * anchor * off[2] : 0);
                    if (c->decomp) c->data[idx] = c->unpred_in[c->unpred_pos++];
                    else FN(push_unpred)(c, c->data[idx]);
                }
        L--;
    }
    for (int level = L; level > 0 && level <= L; level--) {
        double cur = eb;
        if (alpha < 0) {
            cur = (level >= 3) ? eb * 0.5 : eb;
        } else if (alpha >= 1) {
            double ratio = pow(alpha, level - 1);
            if (ratio > beta) ratio = beta;
            cur = eb / ratio;
        }
        c->eb = cur;
        c->ebx2 = 2 * cur;
        c->recip = 1.0 / c->ebx2;
        c->ebf = (float)cur;
        size_t st = (size_t)1 << (level - 1);
        if (FN(level)(c, D, N, cubic, dir, st, level == 1)) return -1;
    }
    return 0;
}

#undef FN
#undef CAT
#undef CAT2
