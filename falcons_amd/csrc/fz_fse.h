// FSE (finite-state entropy, zstd 1.4 / FSE 0.9.0 as vendored by Falcons under
// include/Falcons/encoder/zfse) table construction, usable from host and device code.
//
// The bytes these routines produce must match the vendored library exactly, so each function keeps
// the published FSE steps: normalization (fse_compress.c:435-481 and the fallback :347-432), the
// table header (fse_compress.c:193-275), the symbol spread and state tables (fse_compress.c:65-155,
// fse_decompress.c:70-127) and the header reader (entropy_common.c:41-140).
#pragma once
#include <stdint.h>

#ifndef FZ_HD
#define FZ_HD __host__ __device__ inline
#endif

namespace fz {

constexpr unsigned kFseMaxTableLog = 13;  // zfse/fse.h:641-669 (FSE_MAX_MEMORY_USAGE 15)
constexpr unsigned kFseMinTableLog = 5;
constexpr unsigned kFseDefaultTableLog = 11;
constexpr uint64_t kFseErr = ~0ull;

FZ_HD unsigned highbit32(uint32_t v) { return 31u - (unsigned)__builtin_clz(v); }

// Encoding table of one alphabet (at most 256 symbols, tableLog <= MaxLog).
template <unsigned MaxLog>
struct CTableT {
    uint32_t tableLog;
    uint32_t maxSym;
    uint32_t deltaNbBits[256];
    int32_t deltaFindState[256];
    uint16_t stateTable[1u << MaxLog];
};
using CTable = CTableT<kFseMaxTableLog>;

// Decoding table entry: next-state base, symbol, bits to read.
struct alignas(4) DEntry {
    uint16_t newState;
    uint8_t symbol;
    uint8_t nbBits;
};

FZ_HD unsigned min_table_log(uint64_t srcSize, unsigned maxSym) {
    unsigned a = highbit32((uint32_t)srcSize) + 1;
    unsigned b = highbit32(maxSym) + 2;
    return a < b ? a : b;
}

FZ_HD unsigned optimal_table_log(unsigned maxTableLog, uint64_t srcSize, unsigned maxSym) {
    uint32_t maxBitsSrc = highbit32((uint32_t)(srcSize - 1)) - 2u;
    uint32_t tl = maxTableLog ? maxTableLog : kFseDefaultTableLog;
    uint32_t minBits = min_table_log(srcSize, maxSym);
    if (maxBitsSrc < tl) tl = maxBitsSrc;
    if (minBits > tl) tl = minBits;
    if (tl < kFseMinTableLog) tl = kFseMinTableLog;
    if (tl > kFseMaxTableLog) tl = kFseMaxTableLog;
    return tl;
}

FZ_HD uint64_t normalize_m2(int16_t* norm, uint32_t tableLog, const uint32_t* count, uint64_t total,
                            uint32_t maxSym) {
    const int16_t unset = -2;
    uint32_t distributed = 0;
    const uint32_t lowThreshold = (uint32_t)(total >> tableLog);
    uint32_t lowOne = (uint32_t)((total * 3) >> (tableLog + 1));
    for (uint32_t s = 0; s <= maxSym; s++) {
        if (count[s] == 0) { norm[s] = 0; continue; }
        if (count[s] <= lowThreshold) { norm[s] = -1; distributed++; total -= count[s]; continue; }
        if (count[s] <= lowOne) { norm[s] = 1; distributed++; total -= count[s]; continue; }
        norm[s] = unset;
    }
    uint32_t toDistribute = (1u << tableLog) - distributed;
    if (toDistribute == 0) return 0;
    if ((total / toDistribute) > lowOne) {
        lowOne = (uint32_t)((total * 3) / (toDistribute * 2));
        for (uint32_t s = 0; s <= maxSym; s++)
            if (norm[s] == unset && count[s] <= lowOne) { norm[s] = 1; distributed++; total -= count[s]; }
        toDistribute = (1u << tableLog) - distributed;
    }
    if (distributed == maxSym + 1) {
        uint32_t maxV = 0, maxC = 0;
        for (uint32_t s = 0; s <= maxSym; s++)
            if (count[s] > maxC) { maxV = s; maxC = count[s]; }
        norm[maxV] = (int16_t)(norm[maxV] + (int16_t)toDistribute);
        return 0;
    }
    if (total == 0) {
        for (uint32_t s = 0; toDistribute > 0; s = (s + 1) % (maxSym + 1))
            if (norm[s] > 0) { toDistribute--; norm[s]++; }
        return 0;
    }
    const uint64_t vStepLog = 62 - tableLog;
    const uint64_t mid = (1ull << (vStepLog - 1)) - 1;
    const uint64_t rStep = (((uint64_t)1 << vStepLog) * toDistribute + mid) / total;
    uint64_t tmpTotal = mid;
    for (uint32_t s = 0; s <= maxSym; s++) {
        if (norm[s] == unset) {
            uint64_t end = tmpTotal + (count[s] * rStep);
            uint32_t w = (uint32_t)(end >> vStepLog) - (uint32_t)(tmpTotal >> vStepLog);
            if (w < 1) return kFseErr;
            norm[s] = (int16_t)w;
            tmpTotal = end;
        }
    }
    return 0;
}

// Returns tableLog, 0 when one symbol holds every count, or kFseErr.
FZ_HD uint64_t normalize_count(int16_t* norm, unsigned tableLog, const uint32_t* count, uint64_t total,
                               unsigned maxSym) {
    if (tableLog == 0) tableLog = kFseDefaultTableLog;
    if (tableLog < kFseMinTableLog || tableLog > kFseMaxTableLog) return kFseErr;
    if (tableLog < min_table_log(total, maxSym)) return kFseErr;
    const uint32_t rtb[8] = {0, 473195, 504333, 520860, 550000, 700000, 750000, 830000};
    const uint64_t scale = 62 - tableLog;
    const uint64_t step = ((uint64_t)1 << 62) / total;
    const uint64_t vStep = 1ull << (scale - 20);
    int still = 1 << tableLog;
    unsigned largest = 0;
    int16_t largestP = 0;
    const uint32_t lowThreshold = (uint32_t)(total >> tableLog);
    for (unsigned s = 0; s <= maxSym; s++) {
        if (count[s] == total) return 0;
        if (count[s] == 0) { norm[s] = 0; continue; }
        if (count[s] <= lowThreshold) {
            norm[s] = -1;
            still--;
        } else {
            const uint64_t scaled = (uint64_t)count[s] * step;
            int16_t p = (int16_t)(scaled >> scale);
            if (p < 8) {
                const uint64_t restToBeat = vStep * rtb[p];
                p = (int16_t)(p + ((scaled - ((uint64_t)p << scale)) > restToBeat));
            }
            if (p > largestP) { largestP = p; largest = s; }
            norm[s] = p;
            still -= p;
        }
    }
    if (-still >= (norm[largest] >> 1)) {
        if (normalize_m2(norm, tableLog, count, total, maxSym) == kFseErr) return kFseErr;
    } else {
        norm[largest] = (int16_t)(norm[largest] + still);
    }
    return tableLog;
}

// Table header; returns its byte length or kFseErr. out needs room for (256*14)/8 + 8 bytes.
FZ_HD uint64_t write_ncount(uint8_t* out0, const int16_t* norm, unsigned maxSym, unsigned tableLog) {
    uint8_t* out = out0;
    const int tableSize = 1 << tableLog;
    uint32_t bits = (uint32_t)(tableLog - kFseMinTableLog);
    int nbits = 4;
    unsigned sym = 0;
    const unsigned alphabet = maxSym + 1;
    bool prev0 = false;
    int remaining = tableSize + 1, threshold = tableSize, nbBits = (int)tableLog + 1;
    while (sym < alphabet && remaining > 1) {
        if (prev0) {
            unsigned start = sym;
            while (sym < alphabet && !norm[sym]) sym++;
            if (sym == alphabet) break;
            while (sym >= start + 24) {
                start += 24;
                bits += 0xFFFFu << nbits;
                out[0] = (uint8_t)bits; out[1] = (uint8_t)(bits >> 8); out += 2;
                bits >>= 16;
            }
            while (sym >= start + 3) { start += 3; bits += 3u << nbits; nbits += 2; }
            bits += (uint32_t)(sym - start) << nbits;
            nbits += 2;
            if (nbits > 16) {
                out[0] = (uint8_t)bits; out[1] = (uint8_t)(bits >> 8); out += 2;
                bits >>= 16; nbits -= 16;
            }
        }
        int c = norm[sym++];
        const int max = (2 * threshold - 1) - remaining;
        remaining -= c < 0 ? -c : c;
        c++;
        if (c >= threshold) c += max;
        bits += (uint32_t)c << nbits;
        nbits += nbBits;
        nbits -= (c < max);
        prev0 = (c == 1);
        if (remaining < 1) return kFseErr;
        while (remaining < threshold) { nbBits--; threshold >>= 1; }
        if (nbits > 16) {
            out[0] = (uint8_t)bits; out[1] = (uint8_t)(bits >> 8); out += 2;
            bits >>= 16; nbits -= 16;
        }
    }
    if (remaining != 1) return kFseErr;
    out[0] = (uint8_t)bits;
    out[1] = (uint8_t)(bits >> 8);
    out += (nbits + 7) / 8;
    return (uint64_t)(out - out0);
}

// Little-endian bit peek used by the header reader; bits past `size` read as zero.
FZ_HD uint32_t peek_le_bits(const uint8_t* p, uint64_t size, uint64_t bit, unsigned nb) {
    uint32_t v = 0;
    for (unsigned j = 0; j < nb; j++) {
        uint64_t b = bit + j;
        uint32_t x = (b / 8 < size) ? ((p[b / 8] >> (b % 8)) & 1u) : 0u;
        v |= x << j;
    }
    return v;
}

// Reads a table header; returns bytes consumed or kFseErr. maxSym is in/out.
FZ_HD uint64_t read_ncount(int16_t* norm, unsigned* maxSym, unsigned* tableLog, const uint8_t* src, uint64_t size) {
    const unsigned maxSV = *maxSym;
    for (unsigned s = 0; s <= maxSV; s++) norm[s] = 0;
    uint64_t bit = 0;
    const unsigned tl = peek_le_bits(src, size, bit, 4) + kFseMinTableLog;
    bit += 4;
    if (tl > kFseMaxTableLog) return kFseErr;
    *tableLog = tl;
    int remaining = (1 << tl) + 1, threshold = 1 << tl, nbBits = (int)tl + 1;
    unsigned ch = 0;
    bool prev0 = false;
    while (remaining > 1 && ch <= maxSV) {
        if (prev0) {
            unsigned n0 = ch;
            while (peek_le_bits(src, size, bit, 16) == 0xFFFF) { n0 += 24; bit += 16; }
            while (peek_le_bits(src, size, bit, 2) == 3) { n0 += 3; bit += 2; }
            n0 += peek_le_bits(src, size, bit, 2);
            bit += 2;
            if (n0 > maxSV) return kFseErr;
            while (ch < n0) norm[ch++] = 0;
        }
        const int max = (2 * threshold - 1) - remaining;
        const uint32_t v = peek_le_bits(src, size, bit, (unsigned)nbBits);
        int c;
        if ((int)(v & (uint32_t)(threshold - 1)) < max) {
            c = (int)(v & (uint32_t)(threshold - 1));
            bit += (uint64_t)(nbBits - 1);
        } else {
            c = (int)(v & (uint32_t)(2 * threshold - 1));
            if (c >= threshold) c -= max;
            bit += (uint64_t)nbBits;
        }
        c--;
        remaining -= c < 0 ? -c : c;
        norm[ch++] = (int16_t)c;
        prev0 = (c == 0);
        while (remaining < threshold) { nbBits--; threshold >>= 1; }
    }
    if (remaining != 1) return kFseErr;
    *maxSym = ch - 1;
    return (bit + 7) / 8;
}

// Symbol spread shared by the encoder and decoder tables; tableSymbol must hold 1<<tableLog bytes.
// Returns false if the counts do not fill the table exactly.
FZ_HD bool spread_symbols(uint8_t* tableSymbol, const int16_t* norm, unsigned maxSym, unsigned tableLog,
                          uint32_t* highOut) {
    const uint32_t tableSize = 1u << tableLog, mask = tableSize - 1;
    const uint32_t step = (tableSize >> 1) + (tableSize >> 3) + 3;
    uint32_t high = tableSize - 1;
    for (uint32_t s = 0; s <= maxSym; s++)
        if (norm[s] == -1) tableSymbol[high--] = (uint8_t)s;
    uint32_t pos = 0;
    for (uint32_t s = 0; s <= maxSym; s++)
        for (int k = 0; k < norm[s]; k++) {
            tableSymbol[pos] = (uint8_t)s;
            pos = (pos + step) & mask;
            while (pos > high) pos = (pos + step) & mask;
        }
    if (highOut) *highOut = high;
    return pos == 0;
}

FZ_HD void build_ctable(CTable* ct, const int16_t* norm, unsigned maxSym, unsigned tableLog, uint8_t* scratch) {
    const uint32_t tableSize = 1u << tableLog;
    ct->tableLog = tableLog;
    ct->maxSym = maxSym;
    spread_symbols(scratch, norm, maxSym, tableLog, nullptr);
    uint32_t cumul[257];
    cumul[0] = 0;
    for (uint32_t u = 1; u <= maxSym + 1; u++)
        cumul[u] = cumul[u - 1] + (norm[u - 1] == -1 ? 1u : (uint32_t)norm[u - 1]);
    for (uint32_t u = 0; u < tableSize; u++) {
        const uint8_t s = scratch[u];
        ct->stateTable[cumul[s]++] = (uint16_t)(tableSize + u);
    }
    unsigned total = 0;
    for (unsigned s = 0; s <= maxSym; s++) {
        const int nc = norm[s];
        if (nc == 0) {
            ct->deltaNbBits[s] = ((tableLog + 1) << 16) - (1u << tableLog);
            ct->deltaFindState[s] = 0;
        } else if (nc == -1 || nc == 1) {
            ct->deltaNbBits[s] = (tableLog << 16) - (1u << tableLog);
            ct->deltaFindState[s] = (int32_t)total - 1;
            total++;
        } else {
            const uint32_t maxBitsOut = tableLog - highbit32((uint32_t)(nc - 1));
            ct->deltaNbBits[s] = (maxBitsOut << 16) - ((uint32_t)nc << maxBitsOut);
            ct->deltaFindState[s] = (int32_t)total - nc;
            total += (unsigned)nc;
        }
    }
}

FZ_HD bool build_dtable(DEntry* dt, const int16_t* norm, unsigned maxSym, unsigned tableLog, uint8_t* scratch) {
    const uint32_t tableSize = 1u << tableLog;
    if (!spread_symbols(scratch, norm, maxSym, tableLog, nullptr)) return false;
    uint16_t next[256];
    for (uint32_t s = 0; s <= maxSym; s++) next[s] = norm[s] == -1 ? 1 : (uint16_t)norm[s];
    for (uint32_t u = 0; u < tableSize; u++) {
        const uint8_t s = scratch[u];
        const uint32_t ns = next[s]++;
        const uint8_t nb = (uint8_t)(tableLog - highbit32(ns));
        dt[u].symbol = s;
        dt[u].nbBits = nb;
        dt[u].newState = (uint16_t)((ns << nb) - tableSize);
    }
    return true;
}

// FSE_initCState2 (zfse/fse.h:504-515)
template <class CT>
FZ_HD uint32_t init_state2(const CT* ct, unsigned sym) {
    const uint32_t dnb = ct->deltaNbBits[sym];
    const uint32_t nbOut = (dnb + (1u << 15)) >> 16;
    const uint32_t v = (nbOut << 16) - dnb;
    return ct->stateTable[(v >> nbOut) + ct->deltaFindState[sym]];
}

}  // namespace fz
