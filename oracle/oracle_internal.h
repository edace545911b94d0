/* Internal helpers of the CPU oracle (test infrastructure only; never linked by the product). */
#ifndef FALCONS_ORACLE_INTERNAL_H
#define FALCONS_ORACLE_INTERNAL_H
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define FO_FSE_MAX_TABLELOG 13     /* zfse/fse.h:641-669: FSE_MAX_MEMORY_USAGE 15 -> tableLog 13 */
#define FO_FSE_MIN_TABLELOG 5      /* zfse/fse.h:673 */
#define FO_FSE_DEFAULT_TABLELOG 11 /* zfse/fse.h:645,672 */
#define FO_FSE_ERROR ((size_t)-1)

static inline unsigned fo_highbit32(uint32_t v) { return 31u - (unsigned)__builtin_clz(v); }

/* ---- LSB-first forward bit writer, the layout of zfse/bitstream.h:183-251 ---- */
typedef struct {
    uint8_t *buf;
    size_t cap, pos; /* bytes written */
    uint64_t acc;
    unsigned nacc;
} fo_bitw;

static inline void fo_bitw_init(fo_bitw *w, uint8_t *buf, size_t cap) {
    w->buf = buf; w->cap = cap; w->pos = 0; w->acc = 0; w->nacc = 0;
}
static inline void fo_bitw_add(fo_bitw *w, uint64_t v, unsigned nb) {
    if (nb == 0) return;
    v &= (nb >= 64) ? ~0ull : ((1ull << nb) - 1);
    w->acc |= v << w->nacc;
    w->nacc += nb;
    while (w->nacc >= 8) {
        if (w->pos < w->cap) w->buf[w->pos] = (uint8_t)w->acc;
        w->pos++;
        w->acc >>= 8;
        w->nacc -= 8;
    }
}
/* end mark then the last partial byte (bitstream.h:245-251); returns total bytes */
static inline size_t fo_bitw_close(fo_bitw *w) {
    fo_bitw_add(w, 1, 1);
    if (w->nacc > 0) {
        if (w->pos < w->cap) w->buf[w->pos] = (uint8_t)w->acc;
        w->pos++;
        w->acc = 0; w->nacc = 0;
    }
    return w->pos;
}

/* ---- backward bit reader over such a stream (bitstream.h:262-454 semantics) ---- */
typedef struct {
    const uint8_t *buf;
    int64_t pos; /* bit position of the next unread bit + 1 (reads go downward) */
} fo_bitr;

static inline int fo_bitr_init(fo_bitr *r, const uint8_t *buf, size_t size) {
    r->buf = buf;
    if (size < 1) return -1;
    uint8_t last = buf[size - 1];
    if (last == 0) return -1;
    r->pos = (int64_t)(size - 1) * 8 + (int64_t)fo_highbit32(last); /* bits below the end mark */
    return 0;
}
static inline uint32_t fo_bitr_read(fo_bitr *r, unsigned nb) {
    uint32_t v = 0;
    for (unsigned j = 0; j < nb; j++) {
        int64_t p = r->pos - (int64_t)nb + (int64_t)j;
        uint32_t bit = (p >= 0) ? ((r->buf[p >> 3] >> (p & 7)) & 1u) : 0u;
        v |= bit << j;
    }
    r->pos -= nb;
    return v;
}

/* ---- FSE tables ---- */
typedef struct {
    unsigned tableLog, maxSym;
    uint16_t stateTable[1u << FO_FSE_MAX_TABLELOG];
    int32_t deltaFindState[256];
    uint32_t deltaNbBits[256];
} fo_ctable;

typedef struct {
    uint16_t newState;
    uint8_t symbol;
    uint8_t nbBits;
} fo_dentry;

typedef struct {
    unsigned tableLog;
    fo_dentry e[1u << FO_FSE_MAX_TABLELOG];
} fo_dtable;

unsigned fo_fse_hist(const uint8_t *src, size_t n, unsigned count[256], unsigned *maxSym);
unsigned fo_fse_optimal_tablelog(unsigned maxTableLog, size_t srcSize, unsigned maxSym);
size_t fo_fse_normalize(short *norm, unsigned tableLog, const unsigned *count, size_t total, unsigned maxSym);
size_t fo_fse_write_ncount(uint8_t *out, size_t cap, const short *norm, unsigned maxSym, unsigned tableLog);
size_t fo_fse_read_ncount(short *norm, unsigned *maxSym, unsigned *tableLog, const uint8_t *src, size_t size);
int fo_fse_build_ctable(fo_ctable *ct, const short *norm, unsigned maxSym, unsigned tableLog);
int fo_fse_build_dtable(fo_dtable *dt, const short *norm, unsigned maxSym, unsigned tableLog);
uint32_t fo_fse_init_state2(const fo_ctable *ct, unsigned sym);
void fo_fse_encode_sym(fo_bitw *w, uint32_t *state, const fo_ctable *ct, unsigned sym);
size_t fo_fse_compress(uint8_t *dst, size_t cap, const uint8_t *src, size_t n);
size_t fo_fse_decompress(uint8_t *dst, size_t n, const uint8_t *src, size_t srcSize);

#endif
