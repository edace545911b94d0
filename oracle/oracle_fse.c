/* CPU oracle: restatement of the vendored FSE 0.9.0 (zstd 1.4) used by the reference entropy coder.
 * Test infrastructure only. Every function follows the named reference function step by step,
 * because the produced bytes have to be identical. */
#include "oracle_internal.h"

#include <stdlib.h>

/* hist.c:22-50 (HIST_count_simple) and hist.c:165-183 (HIST_count): counts all byte values and
 * trims maxSym to the largest symbol present; returns the largest count. */
unsigned fo_fse_hist(const uint8_t *src, size_t n, unsigned count[256], unsigned *maxSym) {
    memset(count, 0, 256 * sizeof(unsigned));
    if (n == 0) { *maxSym = 0; return 0; }
    for (size_t i = 0; i < n; i++) count[src[i]]++;
    unsigned m = 255;
    while (!count[m]) m--;
    *maxSym = m;
    unsigned largest = 0;
    for (unsigned s = 0; s <= m; s++)
        if (count[s] > largest) largest = count[s];
    return largest;
}

/* fse_compress.c:315-322 (FSE_minTableLog) */
static unsigned fo_min_tablelog(size_t srcSize, unsigned maxSym) {
    unsigned a = fo_highbit32((uint32_t)srcSize) + 1;
    unsigned b = fo_highbit32(maxSym) + 2;
    return a < b ? a : b;
}

/* fse_compress.c:324-342 (FSE_optimalTableLog_internal with minus = 2) */
unsigned fo_fse_optimal_tablelog(unsigned maxTableLog, size_t srcSize, unsigned maxSym) {
    uint32_t maxBitsSrc = fo_highbit32((uint32_t)(srcSize - 1)) - 2u; /* unsigned wrap kept */
    uint32_t tl = maxTableLog ? maxTableLog : FO_FSE_DEFAULT_TABLELOG;
    uint32_t minBits = fo_min_tablelog(srcSize, maxSym);
    if (maxBitsSrc < tl) tl = maxBitsSrc;
    if (minBits > tl) tl = minBits;
    if (tl < FO_FSE_MIN_TABLELOG) tl = FO_FSE_MIN_TABLELOG;
    if (tl > FO_FSE_MAX_TABLELOG) tl = FO_FSE_MAX_TABLELOG;
    return tl;
}

/* fse_compress.c:347-432 (FSE_normalizeM2): used when the first pass leaves too much to redistribute */
static size_t fo_normalize_m2(short *norm, uint32_t tableLog, const unsigned *count, size_t total, uint32_t maxSym) {
    const short UNSET = -2;
    uint32_t distributed = 0;
    uint32_t lowThreshold = (uint32_t)(total >> tableLog);
    uint32_t lowOne = (uint32_t)((total * 3) >> (tableLog + 1));
    for (uint32_t s = 0; s <= maxSym; s++) {
        if (count[s] == 0) { norm[s] = 0; continue; }
        if (count[s] <= lowThreshold) { norm[s] = -1; distributed++; total -= count[s]; continue; }
        if (count[s] <= lowOne) { norm[s] = 1; distributed++; total -= count[s]; continue; }
        norm[s] = UNSET;
    }
    uint32_t toDistribute = (1u << tableLog) - distributed;
    if (toDistribute == 0) return 0;
    if ((total / toDistribute) > lowOne) {
        lowOne = (uint32_t)((total * 3) / (toDistribute * 2));
        for (uint32_t s = 0; s <= maxSym; s++) {
            if (norm[s] == UNSET && count[s] <= lowOne) {
                norm[s] = 1; distributed++; total -= count[s];
            }
        }
        toDistribute = (1u << tableLog) - distributed;
    }
    if (distributed == maxSym + 1) {
        uint32_t maxV = 0, maxC = 0;
        for (uint32_t s = 0; s <= maxSym; s++)
            if (count[s] > maxC) { maxV = s; maxC = count[s]; }
        norm[maxV] = (short)(norm[maxV] + (short)toDistribute);
        return 0;
    }
    if (total == 0) {
        for (uint32_t s = 0; toDistribute > 0; s = (s + 1) % (maxSym + 1))
            if (norm[s] > 0) { toDistribute--; norm[s]++; }
        return 0;
    }
    {
        uint64_t vStepLog = 62 - tableLog;
        uint64_t mid = (1ull << (vStepLog - 1)) - 1;
        uint64_t rStep = (((uint64_t)1 << vStepLog) * toDistribute + mid) / total;
        uint64_t tmpTotal = mid;
        for (uint32_t s = 0; s <= maxSym; s++) {
            if (norm[s] == UNSET) {
                uint64_t end = tmpTotal + (count[s] * rStep);
                uint32_t sStart = (uint32_t)(tmpTotal >> vStepLog);
                uint32_t sEnd = (uint32_t)(end >> vStepLog);
                uint32_t weight = sEnd - sStart;
                if (weight < 1) return FO_FSE_ERROR;
                norm[s] = (short)weight;
                tmpTotal = end;
            }
        }
    }
    return 0;
}

/* fse_compress.c:435-481 (FSE_normalizeCount). Returns tableLog, 0 for the single-symbol case,
 * or FO_FSE_ERROR. */
size_t fo_fse_normalize(short *norm, unsigned tableLog, const unsigned *count, size_t total, unsigned maxSym) {
    if (tableLog == 0) tableLog = FO_FSE_DEFAULT_TABLELOG;
    if (tableLog < FO_FSE_MIN_TABLELOG) return FO_FSE_ERROR;
    if (tableLog > FO_FSE_MAX_TABLELOG) return FO_FSE_ERROR;
    if (tableLog < fo_min_tablelog(total, maxSym)) return FO_FSE_ERROR;
    static const uint32_t rtb[8] = {0, 473195, 504333, 520860, 550000, 700000, 750000, 830000};
    uint64_t scale = 62 - tableLog;
    uint64_t step = ((uint64_t)1 << 62) / total;
    uint64_t vStep = 1ull << (scale - 20);
    int still = 1 << tableLog;
    unsigned largest = 0;
    short largestP = 0;
    uint32_t lowThreshold = (uint32_t)(total >> tableLog);
    for (unsigned s = 0; s <= maxSym; s++) {
        if (count[s] == total) return 0;
        if (count[s] == 0) { norm[s] = 0; continue; }
        if (count[s] <= lowThreshold) {
            norm[s] = -1;
            still--;
        } else {
            uint64_t scaled = (uint64_t)count[s] * step;
            short proba = (short)(scaled >> scale);
            if (proba < 8) {
                uint64_t restToBeat = vStep * rtb[proba];
                proba = (short)(proba + ((scaled - ((uint64_t)proba << scale)) > restToBeat));
            }
            if (proba > largestP) { largestP = proba; largest = s; }
            norm[s] = proba;
            still -= proba;
        }
    }
    if (-still >= (norm[largest] >> 1)) {
        size_t e = fo_normalize_m2(norm, tableLog, count, total, maxSym);
        if (e == FO_FSE_ERROR) return e;
    } else {
        norm[largest] = (short)(norm[largest] + still);
    }
    return tableLog;
}

/* fse_compress.c:193-275 (FSE_writeNCount_generic): the table header, 16 bits at a time */
size_t fo_fse_write_ncount(uint8_t *out0, size_t cap, const short *norm, unsigned maxSym, unsigned tableLog) {
    (void)cap;
    uint8_t *out = out0;
    const int tableSize = 1 << tableLog;
    uint32_t bits = 0;
    int nbits = 0;
    unsigned symbol = 0;
    const unsigned alphabet = maxSym + 1;
    int prev0 = 0;
    bits += (uint32_t)(tableLog - FO_FSE_MIN_TABLELOG) << nbits;
    nbits += 4;
    int remaining = tableSize + 1;
    int threshold = tableSize;
    int nbBits = (int)tableLog + 1;
    while (symbol < alphabet && remaining > 1) {
        if (prev0) {
            unsigned start = symbol;
            while (symbol < alphabet && !norm[symbol]) symbol++;
            if (symbol == alphabet) break;
            while (symbol >= start + 24) {
                start += 24;
                bits += 0xFFFFu << nbits;
                out[0] = (uint8_t)bits; out[1] = (uint8_t)(bits >> 8); out += 2;
                bits >>= 16;
            }
            while (symbol >= start + 3) {
                start += 3;
                bits += 3u << nbits;
                nbits += 2;
            }
            bits += (uint32_t)(symbol - start) << nbits;
            nbits += 2;
            if (nbits > 16) {
                out[0] = (uint8_t)bits; out[1] = (uint8_t)(bits >> 8); out += 2;
                bits >>= 16;
                nbits -= 16;
            }
        }
        {
            int c = norm[symbol++];
            const int max = (2 * threshold - 1) - remaining;
            remaining -= c < 0 ? -c : c;
            c++;
            if (c >= threshold) c += max;
            bits += (uint32_t)c << nbits;
            nbits += nbBits;
            nbits -= (c < max);
            prev0 = (c == 1);
            if (remaining < 1) return FO_FSE_ERROR;
            while (remaining < threshold) { nbBits--; threshold >>= 1; }
        }
        if (nbits > 16) {
            out[0] = (uint8_t)bits; out[1] = (uint8_t)(bits >> 8); out += 2;
            bits >>= 16;
            nbits -= 16;
        }
    }
    if (remaining != 1) return FO_FSE_ERROR;
    out[0] = (uint8_t)bits;
    out[1] = (uint8_t)(bits >> 8);
    out += (nbits + 7) / 8;
    return (size_t)(out - out0);
}

/* Reader of the header written above (entropy_common.c:41-140 reads the same format). */
typedef struct { const uint8_t *p; size_t n; size_t bit; } fo_lsbr;
static uint32_t fo_lsb_peek(const fo_lsbr *r, unsigned nb) {
    uint32_t v = 0;
    for (unsigned j = 0; j < nb; j++) {
        size_t b = r->bit + j;
        uint32_t bit = (b / 8 < r->n) ? ((r->p[b / 8] >> (b % 8)) & 1u) : 0u;
        v |= bit << j;
    }
    return v;
}
size_t fo_fse_read_ncount(short *norm, unsigned *maxSym, unsigned *tableLog, const uint8_t *src, size_t size) {
    fo_lsbr r = {src, size, 0};
    unsigned maxSV = *maxSym;
    memset(norm, 0, (maxSV + 1) * sizeof(short));
    unsigned tl = fo_lsb_peek(&r, 4) + FO_FSE_MIN_TABLELOG;
    r.bit += 4;
    if (tl > 15) return FO_FSE_ERROR;
    *tableLog = tl;
    int remaining = (1 << tl) + 1;
    int threshold = 1 << tl;
    int nbBits = (int)tl + 1;
    unsigned charnum = 0;
    int prev0 = 0;
    while (remaining > 1 && charnum <= maxSV) {
        if (prev0) {
            unsigned n0 = charnum;
            while (fo_lsb_peek(&r, 16) == 0xFFFF) { n0 += 24; r.bit += 16; }
            while (fo_lsb_peek(&r, 2) == 3) { n0 += 3; r.bit += 2; }
            n0 += fo_lsb_peek(&r, 2);
            r.bit += 2;
            if (n0 > maxSV) return FO_FSE_ERROR;
            while (charnum < n0) norm[charnum++] = 0;
        }
        {
            const int max = (2 * threshold - 1) - remaining;
            uint32_t v = fo_lsb_peek(&r, (unsigned)nbBits);
            int c;
            if ((int)(v & (uint32_t)(threshold - 1)) < max) {
                c = (int)(v & (uint32_t)(threshold - 1));
                r.bit += (size_t)(nbBits - 1);
            } else {
                c = (int)(v & (uint32_t)(2 * threshold - 1));
                if (c >= threshold) c -= max;
                r.bit += (size_t)nbBits;
            }
            c--;
            remaining -= c < 0 ? -c : c;
            norm[charnum++] = (short)c;
            prev0 = !c;
            while (remaining < threshold) { nbBits--; threshold >>= 1; }
        }
    }
    if (remaining != 1) return FO_FSE_ERROR;
    *maxSym = charnum - 1;
    return (r.bit + 7) / 8;
}

/* fse_compress.c:65-155 (FSE_buildCTable_wksp): spread, state table, symbol transforms */
int fo_fse_build_ctable(fo_ctable *ct, const short *norm, unsigned maxSym, unsigned tableLog) {
    const uint32_t tableSize = 1u << tableLog, mask = tableSize - 1;
    const uint32_t step = (tableSize >> 1) + (tableSize >> 3) + 3;
    uint32_t cumul[258];
    uint8_t tableSymbol[1u << FO_FSE_MAX_TABLELOG];
    uint32_t high = tableSize - 1;
    ct->tableLog = tableLog;
    ct->maxSym = maxSym;
    cumul[0] = 0;
    for (uint32_t u = 1; u <= maxSym + 1; u++) {
        if (norm[u - 1] == -1) {
            cumul[u] = cumul[u - 1] + 1;
            tableSymbol[high--] = (uint8_t)(u - 1);
        } else {
            cumul[u] = cumul[u - 1] + (uint32_t)norm[u - 1];
        }
    }
    cumul[maxSym + 1] = tableSize + 1;
    uint32_t pos = 0;
    for (uint32_t s = 0; s <= maxSym; s++) {
        for (int k = 0; k < norm[s]; k++) {
            tableSymbol[pos] = (uint8_t)s;
            pos = (pos + step) & mask;
            while (pos > high) pos = (pos + step) & mask;
        }
    }
    if (pos != 0) return -1;
    for (uint32_t u = 0; u < tableSize; u++) {
        uint8_t s = tableSymbol[u];
        ct->stateTable[cumul[s]++] = (uint16_t)(tableSize + u);
    }
    unsigned total = 0;
    for (unsigned s = 0; s <= maxSym; s++) {
        int nc = norm[s];
        if (nc == 0) {
            ct->deltaNbBits[s] = ((tableLog + 1) << 16) - (1u << tableLog);
            ct->deltaFindState[s] = 0;
        } else if (nc == -1 || nc == 1) {
            ct->deltaNbBits[s] = (tableLog << 16) - (1u << tableLog);
            ct->deltaFindState[s] = (int32_t)total - 1;
            total++;
        } else {
            uint32_t maxBitsOut = tableLog - fo_highbit32((uint32_t)(nc - 1));
            uint32_t minStatePlus = (uint32_t)nc << maxBitsOut;
            ct->deltaNbBits[s] = (maxBitsOut << 16) - minStatePlus;
            ct->deltaFindState[s] = (int32_t)total - nc;
            total += (unsigned)nc;
        }
    }
    return 0;
}

/* zfse/fse.h:504-515 (FSE_initCState2) */
uint32_t fo_fse_init_state2(const fo_ctable *ct, unsigned sym) {
    uint32_t dnb = ct->deltaNbBits[sym];
    uint32_t nbOut = (dnb + (1u << 15)) >> 16;
    uint32_t v = (nbOut << 16) - dnb;
    return ct->stateTable[(v >> nbOut) + ct->deltaFindState[sym]];
}

/* zfse/fse.h:517-524 (FSE_encodeSymbol) */
void fo_fse_encode_sym(fo_bitw *w, uint32_t *state, const fo_ctable *ct, unsigned sym) {
    uint32_t nbOut = (uint32_t)(((uint64_t)*state + ct->deltaNbBits[sym]) >> 16);
    fo_bitw_add(w, *state, nbOut);
    *state = ct->stateTable[(*state >> nbOut) + ct->deltaFindState[sym]];
}

/* fse_decompress.c:70-127 (FSE_buildDTable) */
int fo_fse_build_dtable(fo_dtable *dt, const short *norm, unsigned maxSym, unsigned tableLog) {
    const uint32_t tableSize = 1u << tableLog, mask = tableSize - 1;
    const uint32_t step = (tableSize >> 1) + (tableSize >> 3) + 3;
    uint16_t next[256];
    uint32_t high = tableSize - 1;
    if (maxSym > 255 || tableLog > FO_FSE_MAX_TABLELOG) return -1;
    dt->tableLog = tableLog;
    for (uint32_t s = 0; s <= maxSym; s++) {
        if (norm[s] == -1) {
            dt->e[high--].symbol = (uint8_t)s;
            next[s] = 1;
        } else {
            next[s] = (uint16_t)norm[s];
        }
    }
    uint32_t pos = 0;
    for (uint32_t s = 0; s <= maxSym; s++) {
        for (int k = 0; k < norm[s]; k++) {
            dt->e[pos].symbol = (uint8_t)s;
            pos = (pos + step) & mask;
            while (pos > high) pos = (pos + step) & mask;
        }
    }
    if (pos != 0) return -1;
    for (uint32_t u = 0; u < tableSize; u++) {
        uint8_t s = dt->e[u].symbol;
        uint32_t ns = next[s]++;
        dt->e[u].nbBits = (uint8_t)(tableLog - fo_highbit32(ns));
        dt->e[u].newState = (uint16_t)((ns << dt->e[u].nbBits) - tableSize);
    }
    return 0;
}

/* fse_compress.c:632-694 (FSE_compress -> FSE_compress_wksp -> FSE_compress_usingCTable_generic).
 * Returns 0 (not compressible), 1 (single symbol), the byte count, or FO_FSE_ERROR. */
size_t fo_fse_compress(uint8_t *dst, size_t cap, const uint8_t *src, size_t n) {
    if (n <= 1) return 0;
    unsigned count[256], maxSym = 255;
    unsigned maxCount = fo_fse_hist(src, n, count, &maxSym);
    if (maxCount == n) return 1;
    if (maxCount == 1) return 0;
    if (maxCount < (n >> 7)) return 0;
    unsigned tableLog = fo_fse_optimal_tablelog(FO_FSE_DEFAULT_TABLELOG, n, maxSym);
    short norm[256];
    size_t r = fo_fse_normalize(norm, tableLog, count, n, maxSym);
    if (r == FO_FSE_ERROR) return r;
    uint8_t *hdr = (uint8_t *)malloc(1024);
    size_t nc = fo_fse_write_ncount(hdr, 1024, norm, maxSym, tableLog);
    if (nc == FO_FSE_ERROR) { free(hdr); return nc; }
    fo_ctable *ct = (fo_ctable *)malloc(sizeof(fo_ctable));
    if (fo_fse_build_ctable(ct, norm, maxSym, tableLog)) { free(hdr); free(ct); return FO_FSE_ERROR; }
    /* the bitstream: encoding runs backwards over src with two interleaved states */
    size_t bcap = n + (n >> 7) + 64;
    uint8_t *bs = (uint8_t *)malloc(bcap);
    fo_bitw w;
    fo_bitw_init(&w, bs, bcap);
    size_t ip = n;
    uint32_t s1, s2;
    if (n & 1) {
        s1 = fo_fse_init_state2(ct, src[--ip]);
        s2 = fo_fse_init_state2(ct, src[--ip]);
        fo_fse_encode_sym(&w, &s1, ct, src[--ip]);
    } else {
        s2 = fo_fse_init_state2(ct, src[--ip]);
        s1 = fo_fse_init_state2(ct, src[--ip]);
    }
    if ((n - 2) & 2) {
        fo_fse_encode_sym(&w, &s2, ct, src[--ip]);
        fo_fse_encode_sym(&w, &s1, ct, src[--ip]);
    }
    while (ip > 0) {
        fo_fse_encode_sym(&w, &s2, ct, src[--ip]);
        fo_fse_encode_sym(&w, &s1, ct, src[--ip]);
        fo_fse_encode_sym(&w, &s2, ct, src[--ip]);
        fo_fse_encode_sym(&w, &s1, ct, src[--ip]);
    }
    fo_bitw_add(&w, s2, tableLog);
    fo_bitw_add(&w, s1, tableLog);
    size_t bsz = fo_bitw_close(&w);
    size_t total = nc + bsz;
    size_t ret;
    if (total >= n - 1 || total > cap) {
        ret = 0;
    } else {
        memcpy(dst, hdr, nc);
        memcpy(dst + nc, bs, bsz);
        ret = total;
    }
    free(hdr); free(ct); free(bs);
    return ret;
}

/* fse_decompress.c:254-282 (FSE_decompress): header, table, then n symbols from two states */
size_t fo_fse_decompress(uint8_t *dst, size_t n, const uint8_t *src, size_t srcSize) {
    short norm[256];
    unsigned maxSym = 255, tableLog = 0;
    size_t hl = fo_fse_read_ncount(norm, &maxSym, &tableLog, src, srcSize);
    if (hl == FO_FSE_ERROR || tableLog > FO_FSE_MAX_TABLELOG) return FO_FSE_ERROR;
    fo_dtable *dt = (fo_dtable *)malloc(sizeof(fo_dtable));
    if (fo_fse_build_dtable(dt, norm, maxSym, tableLog)) { free(dt); return FO_FSE_ERROR; }
    fo_bitr r;
    if (fo_bitr_init(&r, src + hl, srcSize - hl)) { free(dt); return FO_FSE_ERROR; }
    uint32_t st[2];
    st[0] = fo_bitr_read(&r, tableLog);
    st[1] = fo_bitr_read(&r, tableLog);
    for (size_t i = 0; i < n; i++) {
        uint32_t *s = &st[i & 1];
        const fo_dentry e = dt->e[*s];
        dst[i] = e.symbol;
        *s = e.newState + fo_bitr_read(&r, e.nbBits);
    }
    free(dt);
    return n;
}
