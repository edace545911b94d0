/* CPU oracle: restatement of the reference entropy coder RleFseEncoder<int16_t>
 * (include/Falcons/encoder/RleFseEncoder.hpp). Test infrastructure only. */
#include <stdlib.h>

#include "falcons_oracle.h"
#include "oracle_internal.h"

#define FO_FBLK 65536           /* RleFseEncoder.hpp:75 */
#define FO_O1_NCTX 15           /* :344 */
#define FO_O1_PB 13             /* :345 */
#define FO_O1_M (1u << FO_O1_PB)
#define FO_O1_NSYM 256          /* :347, symbol 255 escapes to the side stream */
#define FO_O1_K 4               /* :479 */
#define FO_O1_SEG ((size_t)8 << 20) /* :577 */

/* ---- order-1 lookup tables (RleFseEncoder.hpp:354-391) ---- */
static uint8_t fo_ctx_lut[65536], fo_sym_lut[65536];
static int fo_lut_ready = 0;
static void fo_build_luts(void) {
    if (fo_lut_ready) return;
    const int B = (FO_O1_NCTX - 1) / 2, Bf = (B + 1) / 2; /* 7 magnitude buckets per sign, 4 exact */
    for (int u = 0; u < 65536; u++) {
        int code = (int16_t)u;
        int sgn = code < 0;
        int a = sgn ? -code : code;
        int ctx;
        if (a == 0) {
            ctx = 0;
        } else {
            int m;
            if (a <= Bf) {
                m = a - 1;
            } else {
                m = Bf - 1;
                int v = Bf;
                while (v < a && m < B - 1) { v += (v >> 1) + 1; m++; }
            }
            if (m >= B) m = B - 1;
            ctx = 1 + m * 2 + sgn;
        }
        fo_ctx_lut[u] = (uint8_t)ctx;
        uint32_t zz = ((uint32_t)code << 1) ^ (uint32_t)(code >> 31);
        fo_sym_lut[u] = (uint8_t)(zz < FO_O1_NSYM - 1 ? zz : FO_O1_NSYM - 1);
    }
    fo_lut_ready = 1;
}
int fo_o1_ctx(int code) { fo_build_luts(); return fo_ctx_lut[(uint16_t)code]; }
int fo_o1_sym(int code) { fo_build_luts(); return fo_sym_lut[(uint16_t)code]; }

static inline void put_u64(uint8_t **p, uint64_t v) { memcpy(*p, &v, 8); *p += 8; }
static inline uint64_t get_u64(const uint8_t **p) { uint64_t v; memcpy(&v, *p, 8); *p += 8; return v; }

/* ---- class bytes in 64 K blocks (RleFseEncoder.hpp:150-187) ---- */
static size_t fo_fse_enc_blocks(const uint8_t *src, size_t n, uint8_t *out) {
    size_t p = 0;
    uint8_t *tmp = (uint8_t *)malloc(FO_FBLK + 1024);
    for (size_t off = 0; off < n; off += FO_FBLK) {
        size_t len = n - off < FO_FBLK ? n - off : FO_FBLK;
        size_t r = fo_fse_compress(tmp, FO_FBLK + 1024, src + off, len);
        if (r == FO_FSE_ERROR || r == 0 || r >= len) {
            out[p++] = 0;
            memcpy(out + p, src + off, len);
            p += len;
        } else if (r == 1) {
            out[p++] = 2;
            out[p++] = src[off];
        } else {
            out[p++] = 1;
            uint32_t rr = (uint32_t)r;
            memcpy(out + p, &rr, 4);
            p += 4;
            memcpy(out + p, tmp, r);
            p += r;
        }
    }
    free(tmp);
    return p;
}
static int fo_fse_dec_blocks(const uint8_t *in, size_t n, uint8_t *dst, size_t *consumed) {
    size_t dp = 0, rp = 0;
    while (dp < n) {
        size_t len = n - dp < FO_FBLK ? n - dp : FO_FBLK;
        uint8_t flag = in[rp++];
        if (flag == 0) {
            memcpy(dst + dp, in + rp, len);
            rp += len;
        } else if (flag == 2) {
            memset(dst + dp, in[rp++], len);
        } else {
            uint32_t rr;
            memcpy(&rr, in + rp, 4);
            rp += 4;
            if (fo_fse_decompress(dst + dp, len, in + rp, rr) == FO_FSE_ERROR) return -1;
            rp += rr;
        }
        dp += len;
    }
    if (consumed) *consumed = rp;
    return 0;
}

/* ---- MSB-first mantissa packing shared by the gamma streams (RleFseEncoder.hpp:938-966) ---- */
typedef struct { uint8_t *buf; size_t bp; uint64_t acc; int n; } fo_msbw;
static inline void fo_msb_put(fo_msbw *w, uint64_t v, int k) {
    if (k <= 0) return;
    w->acc = (w->acc << k) | (v & ((1ull << k) - 1));
    w->n += k;
    while (w->n >= 8) { w->n -= 8; w->buf[w->bp++] = (uint8_t)(w->acc >> w->n); }
}
static inline void fo_msb_close(fo_msbw *w) {
    if (w->n > 0) w->buf[w->bp++] = (uint8_t)(w->acc << (8 - w->n));
}
typedef struct { const uint8_t *buf; size_t rp; uint64_t acc; int n; } fo_msbr;
static inline uint64_t fo_msb_get(fo_msbr *r, int k) {
    while (r->n < k) { r->acc = (r->acc << 8) | r->buf[r->rp++]; r->n += 8; }
    r->n -= k;
    return (r->acc >> r->n) & ((k >= 64) ? ~0ull : ((1ull << k) - 1));
}
static inline int fo_floor_log2(uint64_t v) { return 63 - __builtin_clzll(v); }

/* ---- RleFseEncoder::preprocess_encode (:850-895) ---- */
void fo_rle_preprocess(const int16_t *bins, size_t n, int *method, int *use_delta) {
    const int16_t mode = 0;
    size_t NW = 256, WL = 32;
    if (n < NW * WL) { WL = n / NW; if (!WL) WL = 1; }
    size_t step = (n > NW * WL) ? (n - WL) / NW : WL;
    if (!step) step = 1;
    size_t z = 0, zd = 0, cnt = 0, cntd = 0;
    for (size_t w = 0; w < NW; w++) {
        size_t off = w * step;
        if (off + WL > n) break;
        int16_t prev = bins[off];
        for (size_t t = 0; t < WL; t++) {
            int16_t x = bins[off + t];
            if (x == mode) z++;
            if (t > 0 && x == prev) zd++;
            prev = x;
        }
        cnt += WL;
        cntd += WL - 1;
    }
    double p0 = cnt ? 1.0 * (double)z / (double)cnt : 1.0;
    double dz = (cnt && cntd) ? (1.0 * (double)zd / (double)cntd - p0) : 0.0;
    *use_delta = dz > 0.005;
    *method = (p0 >= 0.80) ? 0 : 1;
    if (n < (size_t)20 * FO_O1_NCTX * FO_O1_NSYM) *method = 0;
}

/* ---- escape side stream (:483-575) ---- */
static size_t fo_o1_ovf_encode(const int16_t *ovf, size_t n, uint8_t *out) {
    uint8_t *p = out;
    if (n == 0) { *p++ = 0; return 1; }
    uint8_t *vcl = (uint8_t *)malloc(n);
    uint8_t *vman = (uint8_t *)malloc(2 * n + 16);
    fo_msbw vw = {vman, 0, 0, 0};
    for (size_t i = 0; i < n; i++) {
        int32_t f = ovf[i];
        uint32_t zz = ((uint32_t)f << 1) ^ (uint32_t)(f >> 31);
        uint64_t vv = (uint64_t)zz + 1;
        int vk = fo_floor_log2(vv);
        vcl[i] = (uint8_t)vk;
        fo_msb_put(&vw, vv, vk);
    }
    fo_msb_close(&vw);
    uint8_t *ovcl = (uint8_t *)malloc(n + n / FO_FBLK * 8 + 64);
    size_t s_vcl = fo_fse_enc_blocks(vcl, n, ovcl);
    size_t raw = n * 2, cm = s_vcl + vw.bp + 16;
    if (cm >= raw) {
        *p++ = 0;
        memcpy(p, ovf, n * 2);
        p += n * 2;
    } else {
        *p++ = 1;
        put_u64(&p, s_vcl);
        put_u64(&p, vw.bp);
        memcpy(p, ovcl, s_vcl); p += s_vcl;
        memcpy(p, vman, vw.bp); p += vw.bp;
    }
    free(vcl); free(vman); free(ovcl);
    return (size_t)(p - out);
}
static int fo_o1_ovf_decode(const uint8_t **pp, size_t n, int16_t *ovf) {
    const uint8_t *p = *pp;
    uint8_t mode = *p++;
    if (n == 0) { *pp = p; return 0; }
    if (mode == 0) {
        memcpy(ovf, p, n * 2);
        *pp = p + n * 2;
        return 0;
    }
    uint64_t s_vcl = get_u64(&p), vbp = get_u64(&p);
    const uint8_t *ovcl = p;
    p += s_vcl;
    const uint8_t *vman = p;
    p += vbp;
    uint8_t *vcl = (uint8_t *)malloc(n);
    if (fo_fse_dec_blocks(ovcl, n, vcl, NULL)) { free(vcl); return -1; }
    fo_msbr r = {vman, 0, 0, 0};
    for (size_t i = 0; i < n; i++) {
        int vk = vcl[i];
        uint64_t vv = (1ull << vk) | fo_msb_get(&r, vk);
        uint32_t zz = (uint32_t)(vv - 1);
        ovf[i] = (int16_t)(int32_t)((zz >> 1) ^ (~(zz & 1) + 1));
    }
    free(vcl);
    *pp = p;
    return 0;
}

/* ---- o1_normalize (:410-477) ---- */
static void fo_o1_normalize(const uint32_t *cnt, uint16_t *freq) {
    const int nsym = FO_O1_NSYM, M = (int)FO_O1_M;
    uint64_t total = 0;
    int nonzero = 0;
    for (int s = 0; s < nsym; s++) { total += cnt[s]; nonzero += cnt[s] != 0; }
    if (total == 0) {
        for (int s = 0; s < nsym; s++) freq[s] = 0;
        freq[0] = (uint16_t)M;
        return;
    }
    if (nonzero > 1) {
        short norm[FO_O1_NSYM];
        unsigned c32[FO_O1_NSYM];
        for (int s = 0; s < nsym; s++) c32[s] = cnt[s];
        size_t r = fo_fse_normalize(norm, FO_O1_PB, c32, (size_t)total, (unsigned)(nsym - 1));
        if (r != FO_FSE_ERROR) {
            for (int s = 0; s < nsym; s++) freq[s] = (uint16_t)norm[s];
            return;
        }
    }
    int sum = 0, maxf = 0, argmax = 0;
    for (int s = 0; s < nsym; s++) {
        if (cnt[s]) {
            int f = (int)((uint64_t)cnt[s] * (uint64_t)M / total);
            if (f == 0) f = 1;
            freq[s] = (uint16_t)f;
            sum += f;
            if (f > maxf) { maxf = f; argmax = s; }
        } else {
            freq[s] = 0;
        }
    }
    int e = M - sum;
    if (e > 0) {
        freq[argmax] = (uint16_t)(freq[argmax] + e);
    } else if (e < 0) {
        e = -e;
        int take = (maxf - 1 < e) ? (maxf - 1) : e;
        freq[argmax] = (uint16_t)(freq[argmax] - take);
        e -= take;
        while (e > 0) {
            int mi = -1;
            uint32_t mv = 1;
            for (int s = 0; s < nsym; s++)
                if (freq[s] > mv) { mv = freq[s]; mi = s; }
            if (mi < 0) break;
            take = (int)freq[mi] - 1;
            if (take > e) take = e;
            freq[mi] = (uint16_t)(freq[mi] - take);
            e -= take;
        }
    }
}

/* order-1 histogram of one segment (:605-659) */
void fo_o1_histogram(const int16_t *bins, size_t n, uint32_t *cnt, int16_t *ovf, size_t *novf_k) {
    fo_build_luts();
    size_t cs = (n + FO_O1_K - 1) / FO_O1_K;
    size_t novf = 0;
    memset(cnt, 0, sizeof(uint32_t) * FO_O1_NCTX * FO_O1_NSYM);
    if (n < (1u << 20)) {
        for (int k = 0; k < FO_O1_K; k++) {
            size_t lo = cs * k, hi = lo + cs;
            if (hi > n) hi = n;
            uint16_t prev = 0;
            novf_k[k] = 0;
            for (size_t i = lo; i < hi; i++) {
                uint16_t code = (uint16_t)bins[i];
                int c = fo_ctx_lut[prev], s = fo_sym_lut[code];
                if (s == FO_O1_NSYM - 1) { ovf[novf++] = (int16_t)code; novf_k[k]++; }
                cnt[(size_t)c * FO_O1_NSYM + s]++;
                prev = code;
            }
        }
    } else {
        for (int k = 0; k < FO_O1_K; k++) {
            size_t lo = cs * k, hi = lo + cs;
            if (hi > n) hi = n;
            size_t b0 = novf;
            for (size_t i = lo; i < hi; i++)
                if (fo_sym_lut[(uint16_t)bins[i]] == FO_O1_NSYM - 1) ovf[novf++] = bins[i];
            novf_k[k] = novf - b0;
        }
        size_t nwin = 2048, wlen = 512;
        size_t step = n / nwin;
        if (step < wlen) step = wlen;
        for (size_t w = 0; w < n; w += step) {
            int prev = (w > 0) ? bins[w - 1] : 0;
            size_t we = w + wlen;
            if (we > n) we = n;
            for (size_t i = w; i < we; i++) {
                int code = bins[i];
                cnt[(size_t)FO_O1_NSYM * fo_ctx_lut[(uint16_t)prev] + fo_sym_lut[(uint16_t)code]]++;
                prev = code;
            }
        }
        for (size_t t = 0; t < (size_t)FO_O1_NCTX * FO_O1_NSYM; t++) cnt[t]++;
    }
}

/* encode_o1tans_block (:597-751) */
static size_t fo_o1_encode_block(const int16_t *bins, size_t n, uint8_t *out) {
    fo_build_luts();
    uint8_t *p = out;
    size_t cs = (n + FO_O1_K - 1) / FO_O1_K;
    int16_t *ovf = (int16_t *)malloc((n ? n : 1) * 2);
    size_t novf_k[FO_O1_K];
    uint32_t *cnt = (uint32_t *)malloc(sizeof(uint32_t) * FO_O1_NCTX * FO_O1_NSYM);
    fo_o1_histogram(bins, n, cnt, ovf, novf_k);
    size_t novf = 0;
    for (int k = 0; k < FO_O1_K; k++) novf += novf_k[k];
    uint16_t nf[FO_O1_NCTX * FO_O1_NSYM];
    fo_ctable *ct = (fo_ctable *)malloc(sizeof(fo_ctable) * FO_O1_NCTX);
    for (int c = 0; c < FO_O1_NCTX; c++) {
        fo_o1_normalize(cnt + c * FO_O1_NSYM, nf + c * FO_O1_NSYM);
        short normc[FO_O1_NSYM];
        for (int s = 0; s < FO_O1_NSYM; s++) normc[s] = (short)nf[c * FO_O1_NSYM + s];
        fo_fse_build_ctable(&ct[c], normc, FO_O1_NSYM - 1, FO_O1_PB);
    }
    /* each chunk is one tANS chain encoded backwards with a per-symbol table switch */
    uint8_t *buf[FO_O1_K];
    size_t blen[FO_O1_K];
    for (int k = 0; k < FO_O1_K; k++) {
        size_t lo = k * cs, hi = lo + cs;
        if (hi > n) hi = n;
        size_t capk = cs * 2 + 4096;
        buf[k] = (uint8_t *)malloc(capk);
        blen[k] = 0;
        if (hi <= lo) continue;
        fo_bitw w;
        fo_bitw_init(&w, buf[k], capk);
        size_t idx = hi - 1;
        int prevc = (idx > lo) ? bins[idx - 1] : 0;
        uint32_t st = fo_fse_init_state2(&ct[fo_ctx_lut[(uint16_t)prevc]], fo_sym_lut[(uint16_t)bins[idx]]);
        for (size_t i = hi - 1; i-- > lo;) {
            int pc = (i > lo) ? bins[i - 1] : 0;
            fo_fse_encode_sym(&w, &st, &ct[fo_ctx_lut[(uint16_t)pc]], fo_sym_lut[(uint16_t)bins[i]]);
        }
        fo_bitw_add(&w, st, FO_O1_PB);
        blen[k] = fo_bitw_close(&w);
    }
    put_u64(&p, n);
    put_u64(&p, FO_O1_K);
    for (int k = 0; k < FO_O1_K; k++) put_u64(&p, blen[k]);
    for (int k = 0; k < FO_O1_K; k++) put_u64(&p, novf_k[k]);
    memcpy(p, nf, sizeof(nf));
    p += sizeof(nf);
    for (int k = 0; k < FO_O1_K; k++) { memcpy(p, buf[k], blen[k]); p += blen[k]; free(buf[k]); }
    p += fo_o1_ovf_encode(ovf, novf, p);
    free(ovf); free(cnt); free(ct);
    return (size_t)(p - out);
}

/* decode_o1tans_block (:774-843) */
static int fo_o1_decode_block(const uint8_t **pp, size_t n, int16_t *out) {
    fo_build_luts();
    const uint8_t *p = *pp;
    (void)get_u64(&p);
    int K = (int)get_u64(&p);
    if (n == 0) { *pp = p; return 0; }
    if (K < 1 || K > 16) return -1;
    size_t cs = (n + K - 1) / K;
    uint64_t blen[16], nvk[16];
    for (int k = 0; k < K; k++) blen[k] = get_u64(&p);
    for (int k = 0; k < K; k++) nvk[k] = get_u64(&p);
    uint16_t nf[FO_O1_NCTX * FO_O1_NSYM];
    memcpy(nf, p, sizeof(nf));
    p += sizeof(nf);
    const uint8_t *sbase = p;
    size_t tot = 0, totov = 0;
    for (int k = 0; k < K; k++) { tot += blen[k]; totov += nvk[k]; }
    p += tot;
    int16_t *ovf = (int16_t *)malloc((totov ? totov : 1) * 2);
    if (fo_o1_ovf_decode(&p, totov, ovf)) { free(ovf); return -1; }
    fo_dtable *dt = (fo_dtable *)malloc(sizeof(fo_dtable) * FO_O1_NCTX);
    for (int c = 0; c < FO_O1_NCTX; c++) {
        short normc[FO_O1_NSYM];
        for (int s = 0; s < FO_O1_NSYM; s++) normc[s] = (short)nf[c * FO_O1_NSYM + s];
        fo_fse_build_dtable(&dt[c], normc, FO_O1_NSYM - 1, FO_O1_PB);
    }
    size_t off = 0, ooff = 0;
    for (int k = 0; k < K; k++) {
        size_t lo = k * cs, hi = lo + cs;
        if (hi > n) hi = n;
        if (hi > lo && blen[k] > 0) {
            fo_bitr r;
            fo_bitr_init(&r, sbase + off, blen[k]);
            uint32_t st = fo_bitr_read(&r, FO_O1_PB);
            int prev = 0;
            size_t ovi = ooff;
            for (size_t i = lo; i < hi; i++) {
                const fo_dentry e = dt[fo_ctx_lut[(uint16_t)prev]].e[st];
                unsigned s = e.symbol;
                st = e.newState + fo_bitr_read(&r, e.nbBits);
                int code = (s == FO_O1_NSYM - 1) ? ovf[ovi++] : (int)((s >> 1) ^ (~(s & 1) + 1));
                out[i] = (int16_t)code;
                prev = code;
            }
        }
        off += blen[k];
        ooff += nvk[k];
    }
    free(dt); free(ovf);
    *pp = p;
    return 0;
}

/* RleFseEncoder::save + ::encode (:897-982, :1072-1075). codes is not modified. */
size_t fo_rle_encode(const int16_t *codes, size_t n, uint8_t *out) {
    int method, use_delta;
    fo_rle_preprocess(codes, n, &method, &use_delta);
    uint8_t *p = out;
    *p++ = (uint8_t)(method | (use_delta ? 0x80 : 0));
    memset(p, 0, 4); /* lo = 0, mode = 0 */
    p += 4;
    int16_t *bins = (int16_t *)malloc((n ? n : 1) * 2);
    memcpy(bins, codes, n * 2);
    if (use_delta)
        for (size_t i = n; i > 1; i--) bins[i - 1] = (int16_t)(bins[i - 1] - bins[i - 2]);
    if (method == 1) {
        size_t S = n / FO_O1_SEG;
        if (S < 1) S = 1;
        if (S > 16) S = 16;
        put_u64(&p, S);
        size_t seg = (n + S - 1) / S;
        for (size_t g = 0; g < S; g++) {
            size_t lo = g * seg, hi = lo + seg;
            if (hi > n) hi = n;
            p += fo_o1_encode_block(bins + lo, hi - lo, p);
        }
        free(bins);
        return (size_t)(p - out);
    }
    uint8_t *rcl = (uint8_t *)malloc(n + 1), *vcl = (uint8_t *)malloc(n + 1);
    uint8_t *rman = (uint8_t *)malloc(n + 16), *vman = (uint8_t *)malloc(2 * n + 16);
    fo_msbw rw = {rman, 0, 0, 0}, vw = {vman, 0, 0, 0};
    size_t nr = 0, nv = 0, i = 0;
    while (i < n) {
        size_t L = 0;
        while (i + L < n && bins[i + L] == 0) L++;
        i += L;
        uint64_t v = (uint64_t)L + 1;
        int k = fo_floor_log2(v);
        rcl[nr++] = (uint8_t)k;
        fo_msb_put(&rw, v, k);
        if (i >= n) break;
        int32_t f = bins[i++];
        uint32_t zz = ((uint32_t)f << 1) ^ (uint32_t)(f >> 31);
        uint64_t vv = (uint64_t)zz + 1;
        int vk = fo_floor_log2(vv);
        vcl[nv++] = (uint8_t)vk;
        fo_msb_put(&vw, vv, vk);
    }
    fo_msb_close(&rw);
    fo_msb_close(&vw);
    uint8_t *orcl = (uint8_t *)malloc(nr + nr / FO_FBLK * 8 + 64);
    uint8_t *ovcl = (uint8_t *)malloc(nv + nv / FO_FBLK * 8 + 64);
    size_t s_rcl = fo_fse_enc_blocks(rcl, nr, orcl), s_vcl = fo_fse_enc_blocks(vcl, nv, ovcl);
    put_u64(&p, nr); put_u64(&p, nv); put_u64(&p, s_rcl); put_u64(&p, s_vcl);
    put_u64(&p, rw.bp); put_u64(&p, vw.bp);
    memcpy(p, orcl, s_rcl); p += s_rcl;
    memcpy(p, ovcl, s_vcl); p += s_vcl;
    memcpy(p, rman, rw.bp); p += rw.bp;
    memcpy(p, vman, vw.bp); p += vw.bp;
    free(rcl); free(vcl); free(rman); free(vman); free(orcl); free(ovcl); free(bins);
    return (size_t)(p - out);
}

/* RleFseEncoder::load + ::decode (:984-1070, :1076-1079) */
int fo_rle_decode(const uint8_t *in, size_t len, size_t n, int16_t *out, size_t *consumed) {
    (void)len;
    const uint8_t *p = in;
    uint8_t m = *p++;
    int use_delta = (m & 0x80) != 0, method = m & 0x7F;
    p += 4;
    if (method == 1) {
        uint64_t S_ = get_u64(&p);
        size_t S = S_ ? (size_t)S_ : 1;
        size_t seg = (n + S - 1) / S;
        for (size_t g = 0; g < S; g++) {
            size_t lo = g * seg, hi = lo + seg;
            if (hi > n) hi = n;
            if (fo_o1_decode_block(&p, hi - lo, out + lo)) return -1;
        }
    } else {
        uint64_t nr = get_u64(&p), nv = get_u64(&p), s_rcl = get_u64(&p), s_vcl = get_u64(&p);
        uint64_t rbp = get_u64(&p), vbp = get_u64(&p);
        const uint8_t *orcl = p; p += s_rcl;
        const uint8_t *ovcl = p; p += s_vcl;
        const uint8_t *rman = p; p += rbp;
        const uint8_t *vman = p; p += vbp;
        uint8_t *drcl = (uint8_t *)malloc(nr ? nr : 1), *dvcl = (uint8_t *)malloc(nv ? nv : 1);
        if (fo_fse_dec_blocks(orcl, nr, drcl, NULL)) return -1;
        if (nv && fo_fse_dec_blocks(ovcl, nv, dvcl, NULL)) return -1;
        fo_msbr rr = {rman, 0, 0, 0}, vr = {vman, 0, 0, 0};
        size_t prod = 0, ri = 0, vi = 0;
        while (prod < n) {
            int k = drcl[ri++];
            uint64_t v = (1ull << k) | fo_msb_get(&rr, k);
            size_t L = (size_t)(v - 1);
            if (L > n - prod) L = n - prod;
            for (size_t j = 0; j < L; j++) out[prod++] = 0;
            if (prod >= n) break;
            int vk = dvcl[vi++];
            uint64_t vv = (1ull << vk) | fo_msb_get(&vr, vk);
            uint32_t zz = (uint32_t)(vv - 1);
            out[prod++] = (int16_t)(int32_t)((zz >> 1) ^ (~(zz & 1) + 1));
        }
        free(drcl); free(dvcl);
    }
    if (use_delta)
        for (size_t i = 1; i < n; i++) out[i] = (int16_t)(out[i - 1] + out[i]);
    if (consumed) *consumed = (size_t)(p - in);
    return 0;
}
