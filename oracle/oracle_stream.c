/* CPU oracle: Config header, dispatcher, generic compressor framing and the zstd stage
 * (include/Falcons/api/falcons.hpp, api/impl/FalconsDispatcher.hpp, api/impl/FalconsImpl.hpp,
 * api/impl/FalconsAlgoInterp.hpp:33-97, compressor/FalconsGenericCompressor.hpp,
 * lossless/Lossless_zstd.hpp, utils/Config.hpp, utils/Statistic.hpp). Test infrastructure only. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle_zstd.h"

#include "falcons_oracle.h"
#include "oracle_internal.h"

#define FO_MAGIC 0xF342F310u   /* version.hpp.in:10 */
#define FO_DATAVER 0x01000000u /* versionInt("1.0.0"), CMakeLists.txt:7 */

static char fo_err[256];
const char *fo_last_error(void) { return fo_err; }
static void set_err(const char *m) { snprintf(fo_err, sizeof fo_err, "%s", m); }

size_t fo_fse_compress_pub(uint8_t *dst, size_t cap, const uint8_t *src, size_t n) { return fo_fse_compress(dst, cap, src, n); }
size_t fo_fse_decompress_pub(uint8_t *dst, size_t n, const uint8_t *src, size_t srcSize) {
    return fo_fse_decompress(dst, n, src, srcSize);
}
size_t fo_fse_normalize_pub(short *norm, unsigned tableLog, const unsigned *count, size_t total, unsigned maxSym) {
    return fo_fse_normalize(norm, tableLog, count, total, maxSym);
}

/* Config(dims...) + setDims (Config.hpp:152-181) and the defaults (Config.hpp:482-511) */
void fo_config_default(fo_config *c, int N, const uint64_t *dims) {
    memset(c, 0, sizeof *c);
    int n = 0;
    uint64_t num = 1;
    for (int i = 0; i < N; i++)
        if (dims[i] > 1) { c->dims[n++] = dims[i]; num *= dims[i]; }
    if (n == 0) { c->dims[0] = 1; n = 1; }
    c->N = n;
    c->num = num;
    c->pred_dim = n;
    c->block_size = (n == 1 ? 128 : (n == 2 ? 16 : 6));
    c->cmpr_algo = 1;
    c->eb_mode = 0;
    c->abs_eb = 1e-3;
    c->lorenzo = 1; c->regression = 1;
    c->quantbin_cnt = 65536;
    c->data_type = 0;
    c->interp_algo = 1;
    c->interp_direction = 0;
    c->interp_anchor_stride = -1;
    c->interp_alpha = 1.25;
    c->interp_beta = 2.0;
}

static inline void w8(uint8_t **p, uint8_t v) { **p = v; (*p)++; }
static inline void w32(uint8_t **p, uint32_t v) { memcpy(*p, &v, 4); *p += 4; }
static inline void w64(uint8_t **p, uint64_t v) { memcpy(*p, &v, 8); *p += 8; }
static inline void wf64(uint8_t **p, double v) { memcpy(*p, &v, 8); *p += 8; }

/* Config::save (Config.hpp:316-362), dims bit-packed LSB-first (ByteUtil.hpp:195-240) */
size_t fo_config_save(const fo_config *c, uint8_t *out) {
    uint8_t *p = out + 1;
    w32(&p, FO_MAGIC);
    w32(&p, FO_DATAVER);
    w8(&p, (uint8_t)(int8_t)c->N);
    uint64_t mx = 0;
    for (int i = 0; i < c->N; i++) if (c->dims[i] > mx) mx = c->dims[i];
    uint8_t bw = 0;
    while (mx > 0) { mx >>= 1; bw++; }
    w8(&p, bw);
    size_t nbits = (size_t)c->N * bw, nbytes = (nbits + 7) / 8;
    memset(p, 0, nbytes);
    for (int i = 0; i < c->N; i++)
        for (int j = 0; j < bw; j++) {
            size_t b = (size_t)i * bw + j;
            if ((c->dims[i] >> j) & 1) p[b / 8] |= (uint8_t)(1u << (b % 8));
        }
    p += nbytes;
    w64(&p, c->num);
    w8(&p, (uint8_t)c->cmpr_algo);
    w8(&p, (uint8_t)c->eb_mode);
    switch (c->eb_mode) {
        case 0: wf64(&p, c->abs_eb); break;
        case 1: wf64(&p, c->rel_eb); break;
        case 2: wf64(&p, c->psnr_eb); break;
        case 3: wf64(&p, c->l2norm_eb); break;
        case 4: case 5: wf64(&p, c->abs_eb); wf64(&p, c->rel_eb); break;
        default: break;
    }
    uint8_t flags = (uint8_t)(((c->lorenzo & 1) << 7) | ((c->lorenzo2 & 1) << 6) | ((c->regression & 1) << 5) |
                              ((c->regression2 & 1) << 4) | ((c->openmp & 1) << 3));
    w8(&p, flags);
    w8(&p, (uint8_t)c->data_type);
    w32(&p, (uint32_t)c->quantbin_cnt);
    w32(&p, (uint32_t)c->block_size);
    w8(&p, (uint8_t)c->pred_dim);
    size_t sz = (size_t)(p - out);
    out[0] = (uint8_t)sz;
    return sz;
}

/* Config::load with its checks (Config.hpp:372-454) plus Falcons_read_config (falcons.hpp:122-141) */
int fo_config_load(fo_config *c, const uint8_t *in, size_t avail, size_t *hdr_len) {
    if (!in || avail < 1) { set_err("Truncated Falcons configuration"); return -1; }
    size_t hs = in[0];
    if (hs < 1 || hs > avail) { set_err("Truncated Falcons configuration"); return -1; }
    const uint8_t *p = in + 1, *end = in + hs;
#define NEED(n) do { if ((size_t)(end - p) < (size_t)(n)) { set_err("Truncated Falcons configuration field"); return -1; } } while (0)
    uint32_t magic, ver;
    NEED(4); memcpy(&magic, p, 4); p += 4;
    NEED(4); memcpy(&ver, p, 4); p += 4;
    NEED(1); int8_t N = (int8_t)*p++;
    NEED(1); uint8_t bw = *p++;
    if (N < 1 || N > 4 || bw < 1 || bw > 64) { set_err("Invalid Falcons dimensions"); return -1; }
    size_t db = ((size_t)N * bw + 7) / 8;
    if ((size_t)(end - p) < db) { set_err("Truncated Falcons dimensions"); return -1; }
    memset(c, 0, sizeof *c);
    c->N = N;
    for (int i = 0; i < N; i++) {
        uint64_t v = 0;
        for (int j = 0; j < bw; j++) {
            size_t b = (size_t)i * bw + j;
            v |= (uint64_t)((p[b / 8] >> (b % 8)) & 1u) << j;
        }
        c->dims[i] = v;
    }
    p += db;
    NEED(8); memcpy(&c->num, p, 8); p += 8;
    uint64_t expected = 1;
    for (int i = 0; i < N; i++) {
        if (c->dims[i] == 0 || c->dims[i] > UINT64_MAX / expected) { set_err("Invalid Falcons element count"); return -1; }
        expected *= c->dims[i];
    }
    if (c->num != expected) { set_err("Falcons dimensions do not match element count"); return -1; }
    NEED(1); c->cmpr_algo = *p++;
    NEED(1); c->eb_mode = *p++;
    if (c->eb_mode > 5 || c->cmpr_algo > 4) { set_err("Invalid Falcons algorithm or error mode"); return -1; }
    switch (c->eb_mode) {
        case 0: NEED(8); memcpy(&c->abs_eb, p, 8); p += 8; break;
        case 1: NEED(8); memcpy(&c->rel_eb, p, 8); p += 8; break;
        case 2: NEED(8); memcpy(&c->psnr_eb, p, 8); p += 8; break;
        case 3: NEED(8); memcpy(&c->l2norm_eb, p, 8); p += 8; break;
        default: NEED(16); memcpy(&c->abs_eb, p, 8); memcpy(&c->rel_eb, p + 8, 8); p += 16; break;
    }
    c->quantbin_cnt = 65536;
    c->lorenzo = 1; c->regression = 1;
    if (p < end) {
        uint8_t f = *p++;
        c->lorenzo = (f >> 7) & 1; c->lorenzo2 = (f >> 6) & 1; c->regression = (f >> 5) & 1;
        c->regression2 = (f >> 4) & 1; c->openmp = (f >> 3) & 1;
    }
    if (p < end) { c->data_type = *p++; }
    if (p < end) { NEED(4); int32_t v; memcpy(&v, p, 4); p += 4; c->quantbin_cnt = v; }
    if (p < end) { NEED(4); int32_t v; memcpy(&v, p, 4); p += 4; c->block_size = v; }
    if (p < end) { NEED(1); c->pred_dim = *p++; }
#undef NEED
    if (magic != FO_MAGIC) { set_err("magic number mismatch, the input data is not compressed by Falcons"); return -1; }
    if (ver != FO_DATAVER) { set_err("Please use a matching Falcons data version"); return -1; }
    if (avail <= hs) { set_err("Missing Falcons compressed payload"); return -1; }
    c->interp_algo = 1; c->interp_anchor_stride = -1; c->interp_alpha = 1.25; c->interp_beta = 2.0;
    *hdr_len = hs;
    return 0;
}

/* data_range (Statistic.hpp:160-204) with openmp off. Fields of 16 values or fewer, and the scalar
 * build, take scalar_minmax (:38-44): std::max / std::min from d[0], so a NaN in d[0] stays and any
 * other NaN is passed over. The AVX2 build takes avx2_minmax (:97-157): L = 32 / sizeof(T) lanes
 * start at d[0]; each full vector updates lane k by MAXPS / MINPS, which return their second
 * operand (the element) unless the first compares greater / less, so a NaN element enters the lane
 * and the next element of that lane replaces it, and with it everything the lane held before; the
 * lanes are then folded into lane 0 with std::max / std::min (a NaN in lane 0 stays, NaNs in the
 * other lanes are passed over) and the tail after the last full vector is scanned with
 * std::max / std::min. The difference is formed in the element type. */
#define FO_MAX(a, b) ((a) < (b) ? (b) : (a)) /* std::max(a, b) */
#define FO_MIN(a, b) ((b) < (a) ? (b) : (a)) /* std::min(a, b) */
#define FO_MAXPS(a, b) ((a) > (b) ? (a) : (b))
#define FO_MINPS(a, b) ((a) < (b) ? (a) : (b))
#define FO_RANGE_BODY(T, L)                                                          \
    const T *d = (const T *)data;                                                    \
    T mn = d[0], mx = d[0];                                                          \
    size_t i = 0;                                                                    \
    if (fo_get_simd() && n > 16) {                                                   \
        T vmx[L], vmn[L];                                                            \
        for (int k = 0; k < L; k++) vmx[k] = vmn[k] = d[0];                          \
        for (; i + L <= n; i += L)                                                   \
            for (int k = 0; k < L; k++) {                                            \
                vmx[k] = FO_MAXPS(vmx[k], d[i + k]);                                 \
                vmn[k] = FO_MINPS(vmn[k], d[i + k]);                                 \
            }                                                                        \
        mx = vmx[0];                                                                 \
        mn = vmn[0];                                                                 \
        for (int k = 1; k < L; k++) { mx = FO_MAX(mx, vmx[k]); mn = FO_MIN(mn, vmn[k]); } \
    } else {                                                                         \
        i = 1;                                                                       \
    }                                                                                \
    for (; i < n; i++) { mx = FO_MAX(mx, d[i]); mn = FO_MIN(mn, d[i]); }

/* Integer fields: scalar_minmax, the difference formed in the element type (wrapping). */
#define FO_IRANGE(T, UT)                                                                          \
    {                                                                                             \
        const T *d = (const T *)data;                                                             \
        T mn = d[0], mx = d[0];                                                                   \
        for (size_t i = 1; i < n; i++) { mx = FO_MAX(mx, d[i]); mn = FO_MIN(mn, d[i]); }          \
        return (double)(T)(UT)((UT)mx - (UT)mn);                                                  \
    }

double fo_data_range(int dtype, const void *data, size_t n) {
    if (dtype == 0) {
        FO_RANGE_BODY(float, 8)
        return (double)(float)(mx - mn);
    }
    if (dtype == 1) {
        FO_RANGE_BODY(double, 4)
        return mx - mn;
    }
    switch (dtype) {
        case 2: FO_IRANGE(uint8_t, uint8_t)
        case 3: FO_IRANGE(int8_t, uint8_t)
        case 4: FO_IRANGE(uint16_t, uint16_t)
        case 5: FO_IRANGE(int16_t, uint16_t)
        case 6: FO_IRANGE(uint32_t, uint32_t)
        case 7: FO_IRANGE(int32_t, uint32_t)
        case 8: FO_IRANGE(uint64_t, uint64_t)
        case 9: FO_IRANGE(int64_t, uint64_t)
        default: return 0;
    }
}
#undef FO_IRANGE
#undef FO_RANGE_BODY

/* calAbsErrorBound (Statistic.hpp:216-240) */
static int fo_abs_eb(fo_config *c, int dtype, const void *data) {
    if (c->eb_mode == 0) return 0;
    double r;
    switch (c->eb_mode) {
        case 1: c->abs_eb = c->rel_eb * fo_data_range(dtype, data, c->num); break;
        case 2: {
            r = fo_data_range(dtype, data, c->num);
            double v1 = c->psnr_eb + 10 * log10(1 - 2.0 / 3.0 * 0.99);
            c->abs_eb = r * pow(10, v1 / (-20));
            break;
        }
        case 3: c->abs_eb = sqrt(3.0 / (double)c->num) * c->l2norm_eb; break;
        case 4: r = c->rel_eb * fo_data_range(dtype, data, c->num); c->abs_eb = c->abs_eb < r ? c->abs_eb : r; break;
        case 5: r = c->rel_eb * fo_data_range(dtype, data, c->num); c->abs_eb = c->abs_eb > r ? c->abs_eb : r; break;
        default: set_err("Error bound mode not supported"); return -1;
    }
    c->eb_mode = 0;
    return 0;
}

/* Lossless_zstd::compress (Lossless_zstd.hpp:29-46). A destination below 8 + ZSTD_compressBound(n)
 * (or below 8 + n for the raw form) is the reference's std::length_error: returns FO_LENGTH_ERROR. */
#define FO_LENGTH_ERROR ((size_t)-1)
static size_t fo_lossless(const uint8_t *src, size_t n, uint8_t *dst, size_t cap) {
    if (cap < 8) return FO_LENGTH_ERROR;
    if (getenv("FALCONS_NOZSTD")) {
        if (cap - 8 < n) return FO_LENGTH_ERROR;
        memcpy(dst, &n, 8);
        memcpy(dst + 8, src, n);
        return n + 8;
    }
    const fo_zstd_api *Z = fo_zstd();
    if (!Z) { set_err("cannot load libzstd"); return 0; }
    if (cap - 8 < Z->compressBound(n)) return FO_LENGTH_ERROR;
    memcpy(dst, &n, 8);
    size_t r = Z->compress(dst + 8, cap - 8, src, n, 3);
    if (Z->isError(r)) { set_err("Zstd compression failed"); return 0; }
    return r + 8;
}

/* Lossless_zstd::decompress (Lossless_zstd.hpp:48-79) */
static uint8_t *fo_unlossless(const uint8_t *src, size_t n, size_t *outlen) {
    if (!src || n < 8) { set_err("Truncated Falcons lossless payload"); return NULL; }
    size_t dl;
    memcpy(&dl, src, 8);
    size_t pl = n - 8;
    int raw = getenv("FALCONS_NOZSTD") != NULL;
    if (raw) {
        if (dl > pl) { set_err("Truncated Falcons raw payload"); return NULL; }
    } else {
        if (!fo_zstd()) { set_err("cannot load libzstd"); return NULL; }
        unsigned long long fs = fo_zstd()->getFrameContentSize(src + 8, pl);
        if (fs == FO_ZSTD_CONTENTSIZE_ERROR || (fs != FO_ZSTD_CONTENTSIZE_UNKNOWN && fs != dl)) {
            set_err("Invalid Falcons Zstd payload size");
            return NULL;
        }
    }
    uint8_t *dst = (uint8_t *)malloc(dl ? dl : 1);
    if (raw) {
        memcpy(dst, src + 8, dl);
    } else {
        size_t r = fo_zstd()->decompress(dst, dl, src + 8, pl);
        if (fo_zstd()->isError(r) || r != dl) { free(dst); set_err("Invalid or truncated Falcons Zstd payload"); return NULL; }
    }
    *outlen = dl;
    return dst;
}

size_t fo_size_bound(const fo_config *c, int dtype) {
    uint8_t tmp[256];
    size_t es = fo_config_save(c, tmp);
    return 4096 + es + (fo_zstd() ? fo_zstd()->compressBound(c->num * fo_elem_size(dtype)) : 0);
}

/* Falcons_compress_Interp + FalconsGenericCompressor::compress: the buffer handed to the lossless
 * stage is [decomposition.save][u64 nCodes][encoder.save][encoder payload] (FGC:61-83). */
static size_t fo_interp_payload(fo_config *c, int dtype, void *data, uint8_t *out, size_t cap, uint8_t *pre,
                                size_t pre_cap, size_t *pre_len) {
    size_t es = fo_elem_size(dtype);
    /* integer elements and 4-D fields take the generic traversal (IDH:2314-2519) */
    const int generic = dtype >= 2 || c->N == 4;
    if (c->interp_anchor_stride < 0) {
        static const int def[4] = {4096, 128, 32, 16};
        c->interp_anchor_stride = def[c->N - 1];
    }
    size_t dims[4];
    for (int i = 0; i < c->N; i++) dims[i] = c->dims[i];
    size_t n = c->num;
    int16_t *codes = (int16_t *)malloc((n + 16) * 2);
    void *unpred = malloc((n + 16) * es);
    size_t nq = 0, nun = 0, anchor = 0;
    int rc = generic ? fo_gen_compress(dtype, c->N, dims, c->interp_algo, c->interp_direction,
                                       c->interp_anchor_stride, c->interp_alpha, c->interp_beta, c->abs_eb, data,
                                       codes, &nq, unpred, n + 16, &nun, &anchor)
                     : fo_interp_compress(dtype, c->N, dims, c->interp_algo, c->interp_direction,
                                          c->interp_anchor_stride, c->interp_alpha, c->interp_beta, c->abs_eb, data,
                                          codes, &nq, unpred, n + 16, &nun, &anchor);
    if (rc) {
        set_err("unsupported interpolation direction");
        free(codes); free(unpred);
        return 0;
    }
    size_t blen = 4096 + nun * es + nq * 3 + 64;
    uint8_t *buf = (uint8_t *)malloc(blen), *p = buf;
    size_t max_dim = 1;
    for (int i = 0; i < c->N; i++) { uint64_t d = dims[i]; memcpy(p, &d, 8); p += 8; if (dims[i] > max_dim) max_dim = dims[i]; }
    w32(&p, (uint32_t)max_dim);
    w32(&p, (uint32_t)c->interp_algo);
    w32(&p, (uint32_t)c->interp_direction);
    w64(&p, anchor);
    wf64(&p, c->interp_alpha);
    wf64(&p, c->interp_beta);
    w8(&p, dtype >= 2 ? 4 : 2); /* LinearQuantizer uid, integral_uid for integers (LinearQuantizer.hpp:142-152, 252-253) */
    wf64(&p, c->abs_eb);
    w32(&p, (uint32_t)(c->quantbin_cnt / 2));
    w32(&p, (uint32_t)nun);
    memcpy(p, unpred, nun * es);
    p += nun * es;
    w64(&p, nq);
    p += fo_rle_encode(codes, nq, p);
    size_t bl = (size_t)(p - buf);
    if (pre && bl <= pre_cap) memcpy(pre, buf, bl);
    if (pre_len) *pre_len = bl;
    size_t r = fo_lossless(buf, bl, out, cap);
    free(buf); free(codes); free(unpred);
    return r;
}

/* Falcons_compress (falcons.hpp:54-97) -> Falcons_compress_dispatcher (FalconsDispatcher.hpp:16-135) */
size_t fo_compress(const fo_config *conf_in, int dtype, void *data, uint8_t *out, size_t cap, uint8_t *pre,
                   size_t pre_cap, size_t *pre_len) {
    fo_config c = *conf_in;
    size_t es = fo_elem_size(dtype);
    if (pre_len) *pre_len = 0;
    if (cap < fo_size_bound(&c, dtype)) { set_err("The buffer for compressed data is not large enough."); return 0; }
    if (c.openmp) { set_err("the OpenMP stream format is not produced by this path"); return 0; }
    uint8_t hdr[256];
    size_t est = fo_config_save(&c, hdr);
    memset(out, 0, est);
    uint8_t *payload = out + est;
    size_t pcap = cap - est;
    if (fo_abs_eb(&c, dtype, data)) return 0;
    size_t plen = 0;
    if (c.abs_eb == 0) c.cmpr_algo = 4;
    if (c.cmpr_algo != 4) {
        if (c.cmpr_algo == 1) {
            if (fo_tune_interp(&c, dtype, data, 8) == 2) {
                set_err("ALGO_INTERP_LORENZO on 1-D data needs the Lorenzo/regression path");
                return 0;
            }
            c.cmpr_algo = 2;
        }
        if (c.cmpr_algo != 2) { set_err("only the interpolation algorithms are on this path"); return 0; }
        plen = fo_interp_payload(&c, dtype, data, payload, pcap, pre, pre_cap, pre_len);
        if (plen == 0) return 0;
        /* the lossless stage did not fit: the dispatcher stores the field, already overwritten with
         * the reconstruction, losslessly instead (FalconsDispatcher.hpp:100-118) */
        if (plen == FO_LENGTH_ERROR) c.cmpr_algo = 4;
    }
    if (c.cmpr_algo == 4) {
        plen = fo_lossless((const uint8_t *)data, c.num * es, payload, pcap);
        if (plen == FO_LENGTH_ERROR) { set_err("The buffer for compressed data is not large enough."); return 0; }
    } else if ((double)(c.num * es) / 1.0 / (double)plen < 3) {
        size_t zc = fo_zstd()->compressBound(c.num * es) + 8;
        uint8_t *z = (uint8_t *)malloc(zc);
        size_t zl = fo_lossless((const uint8_t *)data, c.num * es, z, zc);
        if (zl && zl != FO_LENGTH_ERROR && zl < plen && zl <= pcap) {
            c.cmpr_algo = 4;
            memcpy(payload, z, zl);
            plen = zl;
        }
        free(z);
    }
    size_t cs = fo_config_save(&c, out);
    return cs + plen;
}

/* Falcons_decompress (falcons.hpp:158-180) -> dispatcher -> FalconsGenericCompressor::decompress */
int fo_decompress(const uint8_t *in, size_t len, int dtype, void *out, size_t out_cap_elems, fo_config *conf_out) {
    fo_config c;
    size_t hs;
    if (fo_config_load(&c, in, len, &hs)) return -1;
    if (conf_out) *conf_out = c;
    if (c.num > out_cap_elems) { set_err("The decompression buffer is too small"); return -1; }
    if (c.openmp) { set_err("the OpenMP stream format is not produced by this path"); return -1; }
    size_t es = fo_elem_size(dtype);
    size_t bl;
    uint8_t *buf = fo_unlossless(in + hs, len - hs, &bl);
    if (!buf) return -1;
    if (c.cmpr_algo == 4) {
        if (bl != c.num * es) { free(buf); set_err("Decompressed data size does not match the original data size"); return -1; }
        memcpy(out, buf, bl);
        free(buf);
        return 0;
    }
    if (c.cmpr_algo != 2) { free(buf); set_err("only the interpolation algorithms are on this path"); return -1; }
    const uint8_t *p = buf;
    size_t dims[4];
    for (int i = 0; i < c.N; i++) { uint64_t d; memcpy(&d, p, 8); p += 8; dims[i] = d; }
    uint32_t blocksize; int32_t algo, dir; uint64_t anchor; double alpha, beta, eb; int32_t radius, nun;
    memcpy(&blocksize, p, 4); p += 4;
    memcpy(&algo, p, 4); p += 4;
    memcpy(&dir, p, 4); p += 4;
    memcpy(&anchor, p, 8); p += 8;
    memcpy(&alpha, p, 8); p += 8;
    memcpy(&beta, p, 8); p += 8;
    /* LinearQuantizer::load (LinearQuantizer.hpp:154-170): integers accept both uids, the floating
     * one marking a legacy stream */
    const uint8_t uid = *p++;
    if (!(uid == 2 || (dtype >= 2 && uid == 4))) { free(buf); set_err("LinearQuantizer uid mismatch"); return -1; }
    memcpy(&eb, p, 8); p += 8;
    memcpy(&radius, p, 4); p += 4;
    memcpy(&nun, p, 4); p += 4;
    const uint8_t *unpred = p;
    p += (size_t)nun * es;
    uint64_t nq;
    memcpy(&nq, p, 8); p += 8;
    int16_t *codes = (int16_t *)malloc((nq + 16) * 2);
    if (fo_rle_decode(p, bl - (size_t)(p - buf), nq, codes, NULL)) { free(codes); free(buf); set_err("entropy decode failed"); return -1; }
    void *un = malloc((size_t)nun * es + 8);
    memcpy(un, unpred, (size_t)nun * es);
    int rc = (dtype >= 2 || c.N == 4)
                 ? fo_gen_decompress(dtype, c.N, dims, algo, dir, (long)anchor, alpha, beta, eb, uid == 2, codes, nq, un,
                                     nun, out)
                 : fo_interp_decompress(dtype, c.N, dims, algo, dir, (long)anchor, alpha, beta, eb, codes, nq, un, nun,
                                        out);
    if (rc == -4) set_err("LinearQuantizer integral reconstruction overflow");
    free(un); free(codes); free(buf);
    return rc;
}
